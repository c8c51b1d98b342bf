"""The basis builder on the device (csrc/linalg.hip, cfs_plume_amd/pipeline.py):
the tall-skinny products, the four sparse operators of buildOutOfCoreMatrices
against the oracle's sparse matrices (oracle/products_oracle.py), the SVD of
device snapshot matrices against numpy's SVD of the same snapshots, and the
reduced operators against the oracle's products of the same dense bases."""
import numpy as np
import pytest

from cfs_plume_amd import fluid as FL
from cfs_plume_amd import pipeline as PP
from cfs_plume_amd.hip import DeviceArray
from oracle import products_oracle as PR

pytestmark = pytest.mark.gpu

FULL = (12, 24, 12)
STEPS = 24


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def run(gpu):
    print(f"[case] full-rank snapshots {FULL} steps={STEPS}", flush=True)
    fluid = FL.FluidSolver(FULL)
    snaps = fluid.run(STEPS)
    yield fluid, snaps
    fluid.close()


def test_tall_skinny_products(gpu):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((5, 20001))
    B = rng.standard_normal((7, 20001))
    M = rng.standard_normal((5, 19))
    dA, dB = DeviceArray.from_numpy(A), DeviceArray.from_numpy(B)
    assert rel(PP.gemm_tn(dA, dB), A @ B.T) <= 1e-13
    assert rel(PP.gemm_nn(dA, M).numpy(), M.T @ A) <= 1e-13
    # repeatable: the split-K partials are reduced in a fixed order
    assert np.array_equal(PP.gemm_tn(dA, dB), PP.gemm_tn(dA, dB))


@pytest.mark.parametrize("op", [PP.OP_POISSON, PP.OP_VEL_TO_DIV, PP.OP_PRESSURE_TO_VEL, PP.OP_DAMPING])
def test_operators_match_the_sparse_matrices(gpu, op):
    p = FL.default_params(FULL)
    bc, dx = list(p.boundaries), p.dx
    mats = {PP.OP_POISSON: PR.poisson_flat(FULL), PP.OP_VEL_TO_DIV: PR.vel_to_div(FULL, dx),
            PP.OP_PRESSURE_TO_VEL: PR.pressure_to_vel(FULL, dx), PP.OP_DAMPING: PR.damping_flat(FULL)}
    M = mats[op]
    rng = np.random.default_rng(op)
    V = rng.standard_normal((3, M.shape[1]))
    got = PP.apply_operator(op, FULL, bc, dx, 0.9, DeviceArray.from_numpy(V)).numpy()
    print(f"[ops] operator {op}: rel {rel(got, (M @ V.T).T):.2e}", flush=True)
    assert rel(got, (M @ V.T).T) <= 1e-14


def test_svd_basis_matches_numpy(run):
    _, snaps = run
    S = snaps.numpy("preadvect")
    b = PP.svd_basis(snaps.arrays["preadvect"])
    keep = np.linalg.norm(S, axis=1) > 1e-7
    u, s, _ = np.linalg.svd(S[keep].T, full_matrices=False)
    lam = s ** 2
    k = int(np.sum(lam / lam[0] >= 1e-9))
    U = b.numpy().T
    print(f"[svd] {b.r} modes (numpy {k}) of {b.snapshots} snapshots; pca {b.pca[:3]}", flush=True)
    assert b.r == k and b.snapshots == int(keep.sum())
    assert rel(b.pca[:k], lam[:k]) <= 1e-10
    assert np.max(np.abs(U.T @ U - np.eye(k))) <= 1e-12
    # same subspace, mode by mode up to sign (well-separated values)
    dots = np.abs(np.sum(U * u[:, :k], axis=0))
    sep = np.ones(k, bool)
    sep[1:] &= lam[1:k] / lam[:k - 1] < 0.99
    sep[:-1] &= lam[1:k] / lam[:k - 1] < 0.99
    assert np.all(dots[sep] > 1 - 1e-6), dots[sep]
    assert rel(U @ (U.T @ S[keep].T), S[keep].T) <= 1e-4


def test_reduced_operators_match_the_oracle(run):
    fluid, snaps = run
    bases = {name: PP.svd_basis(snaps.arrays[stream]) for stream, name in PP.STREAM_BASES.items()}
    p, d = snaps.numpy("pressure"), snaps.numpy("divergence")
    both = DeviceArray.from_numpy(np.stack([p, d], axis=1).reshape(2 * STEPS, -1))
    bases["pressure"] = PP.svd_basis(both)
    prm = fluid.params
    got = PP.build_products(FULL, list(prm.boundaries), prm.dx, bases)
    dense = {k: v.numpy().T for k, v in bases.items()}
    want = PR.products(FULL, prm.dx, dense)
    for name in want:
        e = rel(got[name], want[name])
        print(f"[products] {name} {got[name].shape}: rel {e:.2e}", flush=True)
        # inverseProduct goes through inv(A): its error grows with A's condition number
        assert e <= (1e-8 if name == "inverseProduct" else 1e-10), name
