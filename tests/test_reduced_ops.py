"""The reduced per-step matrices of stepPlume
(SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:170-205, 848-854, 942-946)."""
import numpy as np
import pytest

from cfs_plume_amd.matrix_io import read_matrix, write_matrix
from cfs_plume_amd.reduced_ops import ReducedOperators


def _write_set(d, r, rng, A=None):
    mats = {
        "projected.ptof.matrix": rng.standard_normal((r, r)),
        "projected.vtod.matrix": rng.standard_normal((r, r)),
        "projected.ptov.matrix": rng.standard_normal((r, r)),
        "projected.A.matrix": A if A is not None else rng.standard_normal((r, r)) + r * np.eye(r),
        "damping.peeled.matrix": rng.standard_normal((r, r)),
    }
    for name, M in mats.items():
        write_matrix(d / name, M)
    return mats


def test_missing_inverse_product_is_refused(tmp_path):
    """initOutOfCore (:84-139) needs all six files, inverseProduct included."""
    _write_set(tmp_path, 4, np.random.default_rng(3))
    with pytest.raises(FileNotFoundError) as e:
        ReducedOperators.load(str(tmp_path))
    assert "inverseProduct.matrix" in str(e.value)


def test_unreadable_inverse_product_is_rebuilt_and_written(tmp_path):
    r = 5
    rng = np.random.default_rng(1)
    # A = 2 I: A^-1 = I / 2, so inverseProduct = ptov vtod / 2
    mats = _write_set(tmp_path, r, rng, A=2.0 * np.eye(r))
    # present but short: EIGEN::read returns false, and the product is rebuilt
    (tmp_path / "inverseProduct.matrix").write_bytes(b"\x05\x00")
    ops = ReducedOperators.load(str(tmp_path))
    want = 0.5 * mats["projected.ptov.matrix"] @ mats["projected.vtod.matrix"]
    assert np.allclose(ops.inverse_product, want, rtol=0, atol=1e-13)
    assert np.array_equal(read_matrix(tmp_path / "inverseProduct.matrix"), ops.inverse_product)
    # a second load reads the written file
    assert np.array_equal(ReducedOperators.load(str(tmp_path)).inverse_product, ops.inverse_product)


def test_stages_in_step_order(tmp_path):
    r = 4
    rng = np.random.default_rng(2)
    _write_set(tmp_path, r, rng)
    # permutation damping and a projection that keeps only the first entry
    D = np.roll(np.eye(r), 1, axis=0)
    P = np.zeros((r, r))
    P[0, 0] = 1.0
    write_matrix(tmp_path / "damping.peeled.matrix", D)
    write_matrix(tmp_path / "projected.ptof.matrix", P)
    write_matrix(tmp_path / "inverseProduct.matrix", -np.eye(r))
    ops = ReducedOperators.load(str(tmp_path))
    q = np.array([1.0, 2.0, 3.0, 4.0])
    assert np.array_equal(ops.diffuse(q), [4.0, 1.0, 2.0, 3.0])
    # P q + (-I) q
    assert np.array_equal(ops.pressure_project(q), [0.0, -2.0, -3.0, -4.0])
    # diffusion first, then the projection: D then (P - I)
    assert np.array_equal(ops.after_advection(q), [0.0, -1.0, -2.0, -3.0])


def test_missing_inputs_and_bad_shapes_are_refused(tmp_path):
    with pytest.raises(FileNotFoundError):
        ReducedOperators.load(str(tmp_path))
    with pytest.raises(ValueError):
        ReducedOperators(np.eye(3), np.eye(3), np.eye(4))
