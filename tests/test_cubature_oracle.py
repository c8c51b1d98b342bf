"""Cubature reduced advection and its generator on the CPU: the oracle's
restatement pinned by identities, the library's NNLS against scipy's
Lawson-Hanson, the twister against MT19937's published outputs, the
`cubature` file layout, and the generator's training columns and fit."""
import struct

import numpy as np
import pytest
from scipy.optimize import nnls as scipy_nnls

from cfs_plume_amd import _lib, encode_matrix, synth
from cfs_plume_amd import cubature as CB
from oracle import closure_oracle as CO
from oracle import cubature_oracle as QO
from oracle import oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small():
    dims = (10, 9, 8)
    U = synth.synthetic_basis(dims, 4, seed=21)
    ob = O.Basis(buffers=encode_matrix(U, dims).buffers())
    rows = QO.dense_rows(ob, 4)
    yield dims, ob, rows
    ob.close()


def _q(r, seed, scale):
    return scale * synth.gaussian_q(r, seed)


# --- NNLS -------------------------------------------------------------------------------

@pytest.mark.parametrize("m,n,kind", [(20, 10, "gauss"), (10, 20, "gauss"), (200, 150, "gauss"),
                                      (300, 600, "nonneg"), (120, 120, "shifted")])
def test_nnls_matches_scipy(m, n, kind):
    """The library's Lawson-Hanson solver and scipy.optimize.nnls (an
    independent implementation of the same algorithm) agree, and the
    Kuhn-Tucker conditions hold: w = A^T (b - A x) <= 0 where x = 0 and
    w = 0 where x > 0."""
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    b = rng.standard_normal(m)
    if kind == "nonneg":
        A = np.abs(A)
        b = A @ (np.abs(rng.standard_normal(n)) * (rng.random(n) < 0.3))
    elif kind == "shifted":
        A, b = A + 0.5, b + 3.0
    x, rn, iters, capped = CB.nnls(A, b)
    xs, rs = scipy_nnls(A, b, maxiter=50 * n)
    assert not capped
    assert np.all(x >= 0)
    assert abs(rn - rs) <= 1e-9 * max(1.0, rs)
    assert abs(rn - np.linalg.norm(A @ x - b)) <= 1e-9 * max(1.0, rn)
    assert np.max(np.abs(x - xs)) <= 1e-9 * max(1.0, np.max(np.abs(xs)))
    w = A.T @ (b - A @ x)
    scale = np.linalg.norm(A) * np.linalg.norm(b)
    assert np.all(w[x == 0] <= 1e-10 * scale)
    assert np.all(np.abs(w[x > 0]) <= 1e-10 * scale)


def test_nnls_known_answer():
    """A = I, b = (1, -2, 3): x = (1, 0, 3), residual 2; A = [[1, 1]], b = 2:
    the first column enters, x = (2, 0)."""
    x, rn, _, _ = CB.nnls(np.eye(3), np.array([1.0, -2.0, 3.0]))
    assert np.array_equal(x, [1.0, 0.0, 3.0]) and rn == 2.0
    x, rn, _, _ = CB.nnls(np.array([[1.0, 1.0]]), np.array([2.0]))
    assert np.allclose(x, [2.0, 0.0]) and rn < 1e-15


def test_nnls_iteration_cap():
    """At most max_iter_factor * n inner steps; a run that needs no more gives
    the same x under any larger cap, a capped run stops one step past it with
    a non-negative x."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((300, 600))
    b = rng.standard_normal(300)
    x5, _, it5, c5 = CB.nnls(A, b, max_iter_factor=5)
    assert not c5
    x1, _, it1, c1 = CB.nnls(A, b, max_iter_factor=1)
    if it5 > 600:
        assert c1 and it1 == 601 and np.all(x1 >= 0)
    else:
        assert not c1 and it1 == it5 and np.array_equal(x1, x5)


# --- twister ----------------------------------------------------------------------------

def test_mersenne_twister_published_outputs():
    """MT19937 seeded by init_genrand(5489): the first outputs 3499211612,
    581869302, 3890346734 (Matsumoto & Nishimura's reference sequence)."""
    t = CB.MersenneTwister(5489)
    assert [t.rand_int32() for _ in range(3)] == [3499211612, 581869302, 3890346734]
    # randInt(10) masks with 15 and rejects above 10: 3499211612 & 15 = 12 is
    # rejected, 581869302 & 15 = 6 is returned
    t = CB.MersenneTwister(5489)
    assert t.rand_int(10) == 6
    t = CB.MersenneTwister(7)
    draws = [t.rand_int(5) for _ in range(2000)]
    assert min(draws) == 0 and max(draws) == 5
    t = CB.MersenneTwister(5489)
    assert t.rand() == 3499211612 / 4294967295.0


# --- file ---------------------------------------------------------------------------------

def test_cubature_file_layout(tmp_path):
    """writeCubature (CUBATURE_GENERATOR_EIGEN.cpp:50-91): int32 K, int32
    cells[K], FP64 weights[K]; readAdvectionCubature refuses other sizes."""
    import ctypes
    cells = np.array([5, 0, 17], np.int32)
    w = np.array([0.5, 2.25, 1e-3])
    p = tmp_path / "cubature"
    CB.write_cubature(str(p), cells, w)
    assert p.read_bytes() == struct.pack("<i", 3) + cells.tobytes() + w.tobytes()
    lib = _lib.lib()
    h = ctypes.c_void_p()
    (tmp_path / "short").write_bytes(p.read_bytes()[:-1])
    assert lib.cfs_cubature_open_file(str(tmp_path / "short").encode(), None, ctypes.byref(h)) == -3
    assert lib.cfs_cubature_open_file(str(tmp_path / "none").encode(), None, ctypes.byref(h)) == -2
    (tmp_path / "zero").write_bytes(struct.pack("<i", 0))
    assert lib.cfs_cubature_open_file(str(tmp_path / "zero").encode(), None, ctypes.byref(h)) == -3
    assert not h.value


# --- oracle identities ------------------------------------------------------------------

def test_every_cell_with_weight_one_is_the_exact_reduced_advection(small):
    """With every interior cell a key cell of weight 1 and After built from the
    same basis, the cubature sum is U^T A(U q) (closure_oracle.reduced_advect,
    pinned by closed forms): the backtrace, clamp, corner weights and the
    block-ordered accumulation must all be right."""
    dims, ob, rows = small
    n = dims[0] * dims[1] * dims[2]
    cells = np.arange(n)
    after = QO.after_matrices(rows, cells, np.ones(n))
    for seed, scale in ((1, 30.0), (2, 120.0)):  # displacements of about 1 and 4 cells: the clamp matters
        q = _q(ob.r, seed, scale)
        got = QO.cubature_advect(rows, after, cells, dims, q, 1.0)
        want = CO.reduced_advect(ob, q, 1.0)
        assert rel(got, want) <= 1e-12, rel(got, want)


def test_training_columns_sum_to_the_samples(small):
    """getStagedTrainingColumn: the columns of all cells sum to the stacked,
    normalised post-advection samples when those are the exact reduced
    advection of the pre samples (so weights of 1 fit exactly); the library's
    vectorised columns equal the oracle's cell-by-cell ones."""
    dims, ob, rows = small
    n = dims[0] * dims[1] * dims[2]
    pre = np.stack([_q(ob.r, 10 + s, 60.0) for s in range(3)])
    post = np.stack([CO.reduced_advect(ob, q, 1.0) for q in pre])
    mags = QO.magnitudes(post)
    assert np.array_equal(mags, CB.magnitudes(post))
    b = (post * mags[:, None]).reshape(-1)
    total = np.zeros_like(b)
    for i in range(n):
        total += QO.training_column(rows, rows, dims, 1.0, pre, mags, i)
    assert rel(total, b) <= 1e-12
    adv = CB.advected_fields(rows, dims, pre, 1.0)
    for i in (0, 7, n // 2, n - 1):
        want = QO.training_column(rows, rows, dims, 1.0, pre, mags, i)
        got = np.concatenate([(rows[i].T @ adv[s, i]) * mags[s] for s in range(len(pre))])
        assert rel(got, want) <= 1e-13


def test_generator_fits_its_training_data(small):
    """generateImportanceSampledCubature on exact reduced-advection samples:
    non-negative weights on distinct-or-repeated peeled cells, the reported
    relative error is the fit's, the last fit is the NNLS optimum over the
    kept cells (scipy agrees), and the cubature generalises to held-out
    states better than dropping the advection."""
    dims, ob, rows = small
    # 40 samples x r = 4: 160 equations, fitted by at most 6 x 50 candidate cells
    pre = np.stack([_q(ob.r, 100 + s, 60.0) for s in range(40)])
    post = np.stack([CO.reduced_advect(ob, q, 1.0) for q in pre])
    res = CB.generate_cubature(rows, rows, dims, pre, post, 1.0, candidates_per_pass=50, passes=6, tolerance=0.01)
    assert res.passes == len(res.history) and (res.relative_error < 0.01 or res.passes == 6)
    errs_by_pass = [h["rel_error"] for h in res.history]
    assert all(b <= a + 1e-12 for a, b in zip(errs_by_pass, errs_by_pass[1:]))  # each pass refits a superset
    n = dims[0] * dims[1] * dims[2]
    assert len(res.cells) == len(res.weights) > 0
    assert np.all(res.weights > 0) and np.all((res.cells >= 0) & (res.cells < n))
    mags = QO.magnitudes(post)
    b = (post * mags[:, None]).reshape(-1)
    A = np.stack([QO.training_column(rows, rows, dims, 1.0, pre, mags, c) for c in res.cells], axis=1)
    assert abs(rel(A @ res.weights, b) - res.relative_error) <= 1e-9
    xs, _ = scipy_nnls(A, b, maxiter=50 * A.shape[1])
    assert rel(A @ xs, b) >= res.relative_error - 1e-9  # the kept cells' weights are optimal
    print(f"[cubature] {len(res.cells)} key cells after {res.passes} passes, training error {res.relative_error:.3e}")
    after = QO.after_matrices(rows, res.cells, res.weights)
    errs, base = [], []
    for s in range(4):
        q = _q(ob.r, 500 + s, 60.0)
        exact = CO.reduced_advect(ob, q, 1.0)
        cub = QO.cubature_advect(rows, after, res.cells, dims, q, 1.0)
        errs.append(rel(cub, exact))
        base.append(rel(q, exact))
    assert np.mean(errs) < 0.5 * np.mean(base), (errs, base)
