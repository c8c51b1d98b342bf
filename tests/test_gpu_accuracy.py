"""Compressed-versus-uncompressed accuracy harness on the device
(reducedCompressed3DCompareVelocities.cpp:379-433 with compareSubspace,
SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:271-289).

The harness's dense twin is checked against the compressed run with the
compressed basis's own decoded columns as its dense basis: the two runs then
differ only by FP32 rounding (<= 1e-4 rel-L2 per step over 20 steps).  With
the basis the files were encoded from, the error grows as the coding gets
coarser (percent 0.99 against 0.999999).
"""
import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis, ReducedPlume, encode_matrix, synth
from cfs_plume_amd.accuracy import compare_runs
from cfs_plume_amd.closure import PlumeRun
from cfs_plume_amd.matrix_io import read_matrix

pytestmark = pytest.mark.gpu

DIMS = (30, 22, 14)
R = 8
STEPS = 20
RUNS = [PlumeRun(0.0, 0.0, 1.0), PlumeRun(3.0, -2.0, 1.5)]


@pytest.fixture(scope="module")
def U_dense(gpu):
    return synth.synthetic_basis(DIMS, R, seed=3)


def _plume(U_dense, percent):
    gb = CompressedBasis.from_buffers(encode_matrix(U_dense, DIMS, 32, percent, 50).buffers())
    return ReducedPlume(gb, dt0=1.0, alpha=0.99, force=25.0)


def test_dense_twin_of_the_decoded_basis_matches_the_compressed_run(U_dense, tmp_path):
    print(f"[case] accuracy harness dims={DIMS} r={R}: decoded basis as the dense twin", flush=True)
    plume = _plume(U_dense, 0.9995)
    decoded = plume.U.decode_columns().numpy().astype(np.float64)
    errs = compare_runs(plume, decoded, RUNS, STEPS, out_prefix=str(tmp_path / "twin"))
    print(f"[accuracy] decoded twin: max per-step rel-L2 {errs.max():.2e}", flush=True)
    assert errs.shape == (len(RUNS), STEPS)
    assert errs.max() <= 1e-4
    # the trajectories as written: STEPS rows of r
    q_u = read_matrix(tmp_path / "twin.0.qDot.matrix")
    q_c = read_matrix(tmp_path / "twin.0.compressed.qDot.matrix")
    assert q_u.shape == (STEPS, R) and q_c.shape == (STEPS, R)
    assert np.linalg.norm(q_u[-1]) > 0


def test_error_grows_with_coarser_coding(U_dense):
    print(f"[case] accuracy harness dims={DIMS} r={R}: percent 0.999999 vs 0.99", flush=True)
    fine = compare_runs(_plume(U_dense, 0.999999), U_dense, RUNS, STEPS)
    coarse = compare_runs(_plume(U_dense, 0.99), U_dense, RUNS, STEPS)
    print(f"[accuracy] step {STEPS}: percent 0.999999 {fine[:, -1].mean():.2e}, 0.99 {coarse[:, -1].mean():.2e}",
          flush=True)
    assert np.all(np.isfinite(fine)) and np.all(np.isfinite(coarse))
    assert fine[:, -1].mean() < coarse[:, -1].mean()
    assert fine.max() < 1e-2
