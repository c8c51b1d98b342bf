"""Writes the cubature fixture of the 48x64x48, r = 52 plume stand-in.

    python tests/golden/make_cubature.py OUT_DIR      (needs a HIP device)

The synthetic basis (seed 7, percent 0.9995, nBits 32), the reduced plume
(dt0 1, alpha 0.99, force 25) and its seeded runs (seed 2026) are the closure
workload's; the cubature is fitted by the library's generator
(generateImportanceSampledCubature) to the exact reduced advection at 160
states of 120 exact training runs (80 on the trajectories, 80 perturbed by
10 %), 2000 candidates per pass, relative error 1e-3.  Output: OUT_DIR/
cubature_48x64x48_r52 in the reference's `cubature` layout (int32 K, int32
cells[K], FP64 weights[K]) and a JSON note of the fit.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main(out_dir):
    import cfs_plume_amd as C
    from cfs_plume_amd import closure as CL
    from cfs_plume_amd import cubature as CB
    from cfs_plume_amd import synth

    grid, r = (48, 64, 48), 52
    dims = tuple(g - 2 for g in grid)
    bufs = C.encode_matrix(synth.synthetic_basis(dims, r, seed=7), dims, 32, 0.9995, 50, threads=16).buffers()
    U = C.CompressedBasis.from_buffers(bufs)
    plume = C.ReducedPlume(U, dt0=1.0, alpha=0.99, force=25.0)
    runs = C.ReducedPlume.seeded_runs(220, seed=2026)
    b_train = plume.forcing(runs[:200])
    t0 = time.perf_counter()
    states = CL.cubature_training_states(plume, b_train, 120, 100, 160, 0.1)
    res = plume.make_cubature(states, candidates_per_pass=2000, passes=6, tolerance=1e-3,
                              log=lambda m: print(m, flush=True))
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "cubature_48x64x48_r52")
    CB.write_cubature(path, res.cells, res.weights)
    note = {"key_cells": int(len(res.cells)), "passes": res.passes, "training_rel_error": res.relative_error,
            "history": res.history, "fit_s": round(time.perf_counter() - t0, 1)}
    with open(path + ".json", "w") as f:
        json.dump(note, f, indent=1)
    print(json.dumps(note), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.dirname(os.path.abspath(__file__)))
