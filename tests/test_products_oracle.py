"""Pins of the basis builder's operator oracle (oracle/products_oracle.py):
each sparse operator, built with the loops of SUBSPACE_FLUID_3D_EIGEN
(buildFlatA, computeVelocityToDivergence, computePressureToVelocity,
buildPeeledDampingMatrixFlat), equals the full-rank solver's own stage
restated from FLUID_3D_MIC (oracle/fluid_oracle.py: buildA, the divergence
and gradient of project() with its Neumann/wall borders, the full damping
matrix) on random fields (CPU)."""
import numpy as np

from oracle import fluid_oracle as FO
from oracle import products_oracle as PR

FULL = (7, 9, 6)
DX = 1.0 / 9


def peel(f, comps):
    return FO.peel(FULL, f, comps)


def test_poisson_flat_is_the_interior_of_build_a():
    skip = FO.obstacles(FULL)
    A = FO.build_a(FULL, skip)
    X, Y, Z = FULL
    slab = X * Y
    idx = FO._interior(FULL)
    pos = {int(i): k for k, i in enumerate(idx)}
    M = np.zeros((len(idx), len(idx)))
    Ad, Api, Apj, Apk = A
    for k, i in enumerate(idx):
        M[k, k] = Ad[i]
        for j, a in ((i + 1, Api[i]), (i - 1, Api[i - 1]), (i + X, Apj[i]), (i - X, Apj[i - X]),
                     (i + slab, Apk[i]), (i - slab, Apk[i - slab])):
            if int(j) in pos:
                M[k, pos[int(j)]] = a
    assert np.array_equal(PR.poisson_flat(FULL).toarray(), M)


def test_velocity_to_divergence_is_projects_divergence():
    rng = np.random.default_rng(1)
    X, Y, Z = FULL
    v = rng.standard_normal(3 * X * Y * Z)
    skip = FO.obstacles(FULL)
    vb = FO.border(FULL, v, (False, True, False), 2)  # setZeroX, setNeumannY, setZeroZ
    want = peel(FO.divergence(FULL, vb, skip, DX), 1)
    got = PR.vel_to_div(FULL, DX) @ peel(vb, 3)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)


def test_pressure_to_velocity_is_minus_the_gradient():
    rng = np.random.default_rng(2)
    X, Y, Z = FULL
    p = rng.standard_normal(X * Y * Z)
    want = peel(FO.gradient_subtract(FULL, np.zeros(3 * X * Y * Z), p, DX), 3)
    got = PR.pressure_to_vel(FULL, DX) @ peel(p, 1)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


def test_damping_flat_is_the_full_damping():
    rng = np.random.default_rng(3)
    X, Y, Z = FULL
    v = rng.standard_normal(3 * X * Y * Z)
    want = peel(FO.damping(FULL, v), 3)
    got = PR.damping_flat(FULL) @ peel(v, 3)
    assert np.allclose(got, want, rtol=1e-14, atol=1e-15)
