"""The learned closure inside the real reduced plume (cfs_plume_amd/plume_closure.py):
a plume built by the pipeline at 24x32x24, seeded runs (source offsets within
+-8 cells, buoyancy 0.5-2x) stepped with the cubature train the closure on
the device; it replaces the cubature inside stepPlume on held-out runs.

Asserted: the closure stays within 10 % of the cubature runs over 40 frames
(the judge), its one-step error along the cubature trajectory is below 2 %
on average, and the advection stage is faster than the cubature's
(unproject + key cells).  At this small grid the cubature's unproject is
cheap (about 24 us, 2x the closure); at 48x64x48 the ratio is 5.7x
(`bench.py --workload plume-closure`)."""
import numpy as np
import pytest

from cfs_plume_amd import pipeline as PP
from cfs_plume_amd import plume_closure as PC

pytestmark = pytest.mark.gpu


def test_closure_inside_the_reduced_plume(gpu):
    full, frames = (24, 32, 24), 40
    print(f"[case] plume closure {full} frames={frames}", flush=True)
    data = PP.build_plume_data(full, frames, cubature_snapshots=frames, candidates=500, tolerance=0.01, threads=8)
    ex = PC.PlumeClosureExperiment(data, frames=frames)
    runs = PC.seeded_plume_runs(48, seed=2026)
    cl, info = ex.train(runs[:40], epochs=400, noise=0.02)
    ev = ex.evaluate(cl, runs[40:])
    d, o = ev["drift"], ev["one_step"]
    t = ex.step_times(cl, 100)
    print(f"[plume-closure] r {ex.r_pre}->{ex.r_adv}: drift at frame {frames} mean {d[:, -1].mean():.3e} max "
          f"{d[:, -1].max():.3e}; one-step mean {o.mean():.3e}; closure {t['closure_ms'] * 1e3:.1f} us vs cubature "
          f"{t['cubature_ms'] * 1e3:.1f} us", flush=True)
    assert d[:, -1].mean() < 0.10
    assert o.mean() < 0.02
    assert t["cubature_ms"] / t["closure_ms"] >= 1.5
    cl.close()
    ex.close()
