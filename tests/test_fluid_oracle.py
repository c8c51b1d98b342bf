"""Pins of the full-rank plume oracle (oracle/fluid_oracle.py) by what the
mathematics fixes (CPU).

The MIC(0) preconditioner (Bridson's modified incomplete Cholesky,
buildMICPreconditioner FLUID_3D_MIC.cpp:825-862, applied by the two sweeps of
applyMICPreconditioner :864-899) is M = L L^T.  With tau = 0 it is the plain
incomplete Cholesky IC(0), so M equals A on every entry of A's sparsity
pattern; with tau = 1 it is the fully modified one, so M and A have equal
row sums.  Both properties hold for the right indices and signs only, and are
checked on M recovered column by column from the sweeps."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fluid_oracle as FO

FULL = (7, 8, 6)


def interior_matrix(full, A):
    """A restricted to the interior cells (the pressure unknowns), dense."""
    X, Y, Z = full
    slab = X * Y
    Ad, Api, Apj, Apk = A
    idx = FO._interior(full)
    pos = {int(i): k for k, i in enumerate(idx)}
    M = np.zeros((len(idx), len(idx)))
    for k, i in enumerate(idx):
        M[k, k] = Ad[i]
        for j, a in ((i + 1, Api[i]), (i - 1, Api[i - 1]), (i + X, Apj[i]), (i - X, Apj[i - X]),
                     (i + slab, Apk[i]), (i - slab, Apk[i - slab])):
            if int(j) in pos:
                M[k, pos[int(j)]] = a
    return idx, M


def preconditioner_matrix(full, A, pc):
    idx = FO._interior(full)
    cols = []
    for i in idx:
        r = np.zeros(full[0] * full[1] * full[2])
        r[i] = 1.0
        cols.append(FO.apply_mic(full, A, pc, r)[idx])
    return np.linalg.inv(np.array(cols).T)


def test_poisson_matrix_has_neumann_walls_and_open_top():
    skip = FO.obstacles(FULL)
    A = FO.build_a(FULL, skip)
    idx, M = interior_matrix(FULL, A)
    X, Y, Z = FULL
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    # the diagonal counts the non-wall neighbours: walls at x and z, open y
    want = 6.0 - (x == 1) - (x == X - 2) - (z == 1) - (z == Z - 2)
    assert np.array_equal(np.diag(M), want)
    assert np.allclose(M, M.T)
    # rows next to the open top/bottom couple to the border (p = 0 there):
    # their row sums are positive, all others zero
    rs = M.sum(axis=1)
    assert np.all(rs[(y == 1) | (y == Y - 2)] > 0.5)
    assert np.allclose(rs[(y > 1) & (y < Y - 2)], 0.0)


def test_ic0_matches_a_on_its_pattern():
    skip = FO.obstacles(FULL)
    A = FO.build_a(FULL, skip)
    pc = FO.build_precon(FULL, A, tau=0.0)
    _, M_a = interior_matrix(FULL, A)
    M = preconditioner_matrix(FULL, A, pc)
    pattern = M_a != 0.0
    assert np.allclose(M[pattern], M_a[pattern], rtol=1e-12, atol=1e-12)
    assert np.allclose(M, M.T, atol=1e-12)


def test_full_mic_matches_a_row_sums():
    skip = FO.obstacles(FULL)
    A = FO.build_a(FULL, skip)
    pc = FO.build_precon(FULL, A, tau=1.0)
    idx, M_a = interior_matrix(FULL, A)
    M = preconditioner_matrix(FULL, A, pc)
    # The modification moves every fill-in entry of a row onto its diagonal,
    # including fill towards the open top border, which is not an unknown:
    # rows with such fill (a lower neighbour coupled to the border row) are
    # left out; the walls' couplings are zero and cause no fill.
    y = (idx // FULL[0]) % FULL[1]
    rows = y <= FULL[1] - 3
    assert rows.sum() > len(idx) // 2
    assert np.allclose(M.sum(axis=1)[rows], M_a.sum(axis=1)[rows], atol=1e-11)
    assert not np.allclose(M.sum(axis=1), M_a.sum(axis=1), atol=1e-3)


def test_pcg_solves_the_poisson_system():
    skip = FO.obstacles(FULL)
    A = FO.build_a(FULL, skip)
    pc = FO.build_precon(FULL, A)
    idx, M_a = interior_matrix(FULL, A)
    rng = np.random.default_rng(3)
    b = np.zeros(FULL[0] * FULL[1] * FULL[2])
    b[idx] = rng.standard_normal(len(idx))
    x, its, mx = FO.solve_poisson(FULL, A, pc, skip, b)
    want = spla.spsolve(sp.csr_matrix(M_a), b[idx])
    assert 0 < its < 100 and mx <= 1e-10
    assert np.max(np.abs(x[idx] - want)) <= 1e-8 * np.max(np.abs(want))
    assert np.all(x[np.setdiff1d(np.arange(len(b)), idx)] == 0.0)  # borders keep the initial zero


def test_damping_keeps_a_constant_field_inside():
    X, Y, Z = FULL
    v = np.ones(3 * X * Y * Z) * 2.5
    out = FO.damping(FULL, v).reshape(Z, Y, X, 3)
    assert np.allclose(out[2:-2, 2:-2, 2:-2], 2.5, rtol=1e-15)
    # a cell next to the peeled border loses the missing neighbour's share w/6
    assert np.isclose(out[1, 3, 3, 0], 2.5 * (1.0 - 0.9 / 6.0))
    assert np.array_equal(out[0], v.reshape(Z, Y, X, 3)[0])  # the border is untouched


def test_rest_stays_at_rest_and_smoke_rises():
    full = (12, 24, 12)  # the smoke layers y = 1, 2 lie inside
    s = FO.Solver(full)
    n = full[0] * full[1] * full[2]
    v, d, h, snaps, _ = s.step(np.zeros(3 * n), np.zeros(n), np.zeros(n))
    assert not np.any(v) and not np.any(d) and not np.any(snaps["pressure"])
    d, h = s.add_smoke(np.zeros(n), np.zeros(n))
    assert d.sum() > 0
    v, d, h, snaps, (its, mx) = s.step(np.zeros(3 * n), d, h)
    assert its > 0 and mx <= 1e-10
    assert v.reshape(-1, 3)[:, 1].sum() > 0.0  # buoyancy lifts the smoke
    assert set(snaps) == {"preadvect", "prediffuse", "preproject", "final_velocity", "pressure", "divergence"}
