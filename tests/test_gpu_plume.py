"""The whole plume timestep on the device (addSmokeColumn + stepPlume,
SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501) against the FP64 oracle
composed the same way (oracle/plume_oracle.py), with each subspace
advection: exact U^T A(U q), the cubature, the learned closure.  Every frame
is checked from the device's own previous state; velocity, density, heat and
q within 2e-5 rel-L2 (FP32 fields against FP64).  A frame captured in a HIP
graph replays bit for bit."""
import numpy as np
import pytest

from cfs_plume_amd import Closure, CompressedBasis, DeviceArray, Stream, encode_matrix, synth
from cfs_plume_amd import cubature as CB
from cfs_plume_amd import plume as PL
from oracle import closure_oracle as CO
from oracle import cubature_oracle as QO
from oracle import oracle as O
from oracle import plume_oracle as PO

pytestmark = pytest.mark.gpu

TOL = 2e-5
DIMS = (22, 30, 22)
R = 12


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def bases(gpu):
    out = []
    for seed in (31, 32):
        print(f"[case] plume basis dims={DIMS} r={R} seed={seed}", flush=True)
        bufs = encode_matrix(synth.synthetic_basis(DIMS, R, seed=seed), DIMS).buffers()
        out.append((O.Basis(buffers=bufs), CompressedBasis.from_buffers(bufs)))
    rng = np.random.default_rng(5)
    mats = (0.97 * np.eye(R) + 0.01 * rng.standard_normal((R, R)), np.eye(R) + 0.01 * rng.standard_normal((R, R)),
            0.02 * rng.standard_normal((R, R)))
    return out, mats


def _params_dict(p):
    return {"dt": p.dt, "dx": p.dx, "buoyancy": p.buoyancy, "vorticity_eps": p.vorticity_eps,
            "open_x": bool(p.open_x), "open_y": bool(p.open_y), "open_z": bool(p.open_z)}


def _initial_state(full, seed):
    X, Y, Z = full
    rng = np.random.default_rng(seed)
    vel = 0.3 * rng.standard_normal((Z, Y, X, 3))
    dens = rng.random((Z, Y, X))
    heat = rng.random((Z, Y, X))
    return vel, dens, heat


@pytest.mark.parametrize("mode,two_bases", [("exact", False), ("cubature", False), ("closure", False),
                                            ("exact", True)])
def test_plume_frames_match_the_oracle(bases, mode, two_bases):
    (ob_pre, gb_pre), (ob_other, gb_other) = bases[0]
    D, P, I = bases[1]
    ob_fin, gb_fin = (ob_other, gb_other) if two_bases else (ob_pre, gb_pre)
    solver = PL.PlumeSolver(gb_pre, gb_fin, D, P, I)
    prm = _params_dict(solver.params)
    full = solver.full
    X, Y, Z = full
    dt0 = prm["dt"] / prm["dx"]
    keep = None
    if mode == "exact":
        solver.set_advection(PL.ADVECT_EXACT)
        advect = lambda q: CO.reduced_advect(ob_pre, q, dt0)  # noqa: E731
    elif mode == "cubature":
        rng = np.random.default_rng(7)
        n = DIMS[0] * DIMS[1] * DIMS[2]
        cells = np.sort(rng.choice(n, size=400, replace=False)).astype(np.int32)
        w = rng.uniform(0.5, 2.0, size=400)
        keep = CB.Cubature.from_cells(cells, w, gb_pre)
        solver.set_advection(PL.ADVECT_CUBATURE, cubature=keep)
        rows = QO.dense_rows(ob_pre, 8)
        after = QO.after_matrices(rows, cells, w)
        advect = lambda q: QO.cubature_advect(rows, after, cells, DIMS, q, dt0)  # noqa: E731
    else:
        keep = Closure(R, max_batch=16, seed=3)
        rng = np.random.default_rng(9)
        keep.set_scales(rng.uniform(0.5, 2.0, R).astype(np.float32), rng.uniform(0.01, 0.1, R).astype(np.float32))
        solver.set_advection(PL.ADVECT_CLOSURE, closure=keep)
        params = keep.get_params()
        s_in, s_out = keep.get_scales()
        advect = lambda q: CO.closure_step(params, s_in, s_out, q, R)[0]  # noqa: E731
    print(f"[case] plume frames mode={mode} two_bases={two_bases} full={full} dx={prm['dx']:.4f} "
          f"eps={prm['vorticity_eps']}", flush=True)
    vel, dens, heat = _initial_state(full, 11)
    dv = DeviceArray.from_numpy(vel.reshape(-1).astype(np.float32))
    dd = DeviceArray.from_numpy(dens.reshape(-1).astype(np.float32))
    dh = DeviceArray.from_numpy(heat.reshape(-1).astype(np.float32))
    dq = DeviceArray.zeros((R,), np.float64)
    for frame in range(3):
        v0 = dv.numpy().astype(np.float64).reshape(Z, Y, X, 3)
        d0 = dd.numpy().astype(np.float64).reshape(Z, Y, X)
        h0 = dh.numpy().astype(np.float64).reshape(Z, Y, X)
        solver.frame(dv, dd, dh, dq)
        d1, h1 = PO.add_smoke(d0, h0)
        wv, wd, wh, wq = PO.plume_step(ob_pre, ob_fin, D, P, I, prm, v0, d1, h1, advect)
        errs = (rel(dv.numpy(), wv.reshape(-1)), rel(dd.numpy(), wd.reshape(-1)), rel(dh.numpy(), wh.reshape(-1)),
                rel(dq.numpy(), wq))
        print(f"[plume] {mode} frame {frame}: velocity {errs[0]:.2e} density {errs[1]:.2e} heat {errs[2]:.2e} "
              f"q {errs[3]:.2e}", flush=True)
        assert max(errs) <= TOL, errs
    assert gb_pre.check_stream_errors() == 0
    solver.close()


def test_plume_frame_replays_in_a_hip_graph(bases):
    (ob_pre, gb_pre), _ = bases[0]
    D, P, I = bases[1]
    solver = PL.PlumeSolver(gb_pre, gb_pre, D, P, I)
    solver.set_advection(PL.ADVECT_EXACT)
    X, Y, Z = solver.full
    vel, dens, heat = _initial_state(solver.full, 12)
    init = [a.reshape(-1).astype(np.float32) for a in (vel, dens, heat)]
    s = Stream()
    e = [DeviceArray.from_numpy(a) for a in init] + [DeviceArray.zeros((R,), np.float64)]
    solver.frame(*e, stream=s)
    s.synchronize()
    eager = [a.numpy() for a in e]
    g = [DeviceArray.from_numpy(a) for a in init] + [DeviceArray.zeros((R,), np.float64)]
    graph = solver.capture(*g, stream=s)
    for rep in range(3):
        for a, src in zip(g, init + [np.zeros(R)]):
            a.upload(src)
        graph.launch()
        s.synchronize()
        for i, (a, want) in enumerate(zip(g, eager)):
            assert np.array_equal(a.numpy(), want), f"replay {rep}: array {i}"
        solver.frame(*e, stream=s)  # eager work between replays
        s.synchronize()
    solver.close()
