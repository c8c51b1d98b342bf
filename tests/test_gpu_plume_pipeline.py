"""The plume data pipeline end to end on the device (cfs_plume_amd/pipeline.py,
accuracy.compare_plume): a full-rank plume, its SVD bases, the reduced
operators, the cubature from the snapshot pairs, the compressed bases, and
the reduced plume stepped with the compressed and the uncompressed bases.

What must hold: the reduced model with the exact advection reproduces the
full-rank run it was built from (its bases span every snapshot, so the
reduced operators reproduce each stage: relative velocity error well below
1e-2 at every frame); the cubature stays close to it; compression at percent
0.9995 changes q by a few percent at most."""
import numpy as np
import pytest

from cfs_plume_amd import accuracy as A
from cfs_plume_amd import pipeline as PP

pytestmark = pytest.mark.gpu

FULL = (24, 32, 24)
STEPS = 40


def test_pipeline_reproduces_the_full_rank_run(gpu):
    print(f"[case] plume pipeline {FULL} steps={STEPS}", flush=True)
    data = PP.build_plume_data(FULL, STEPS, cubature_snapshots=STEPS, candidates=500, tolerance=0.01, threads=8,
                               log=lambda m: print(m, flush=True))
    ranks = {k: b.r for k, b in data.bases.items()}
    print(f"[pipeline] ranks {ranks}, cubature {len(data.cubature.cells)} cells, error "
          f"{data.cubature.relative_error:.2e}", flush=True)
    assert all(r >= 3 for r in ranks.values())
    assert data.cubature.relative_error < 0.01
    acc = A.compare_plume(data, STEPS)
    ex = acc["uncompressed_exact"]["velocity_vs_full_rank"]
    cu = acc["uncompressed"]["velocity_vs_full_rank"]
    co = acc["compressed"]["velocity_vs_full_rank"]
    q = acc["q_rel_l2"]
    print(f"[pipeline] velocity vs full rank at the last frame: exact {ex[-1]:.2e}, cubature {cu[-1]:.2e}, "
          f"compressed {co[-1]:.2e}; compressed vs uncompressed q: first {q[0]:.2e} max {q.max():.2e}", flush=True)
    assert ex.max() < 1e-2
    assert cu.max() < 0.05
    assert q.max() < 0.05
