"""Pins of the plume-step oracle (oracle/plume_oracle.py) by closed forms (CPU)."""
import numpy as np

from oracle import plume_oracle as PO


def grid(Z, Y, X):
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return z.astype(float), y.astype(float), x.astype(float)


def test_curl_of_a_rigid_rotation():
    """v = Omega (-(y - yc), x - xc, 0) in physical units (cell * dx): the curl
    is (0, 0, 2 Omega) at every interior cell, since central and one-sided
    differences are exact on linear fields; |w| is uniform, so N = 0 and the
    confinement adds nothing (dyadic values keep every difference exact:
    with a rounding-level gradient N would be normalised to a unit vector)."""
    Z, Y, X, dx, om = 8, 10, 9, 0.125, 0.5
    z, y, x = grid(Z, Y, X)
    v = np.stack([-om * (y - 4.5) * dx, om * (x - 3.0) * dx, np.zeros_like(x)], axis=-1)
    w, wmag = PO.curl(v, 1.0 / dx)
    inner = (slice(1, -1),) * 3
    assert np.allclose(w[inner][..., 0], 0, atol=1e-13) and np.allclose(w[inner][..., 1], 0, atol=1e-13)
    assert np.allclose(w[inner][..., 2], 2 * om, rtol=1e-13)
    assert np.all(w[0] == 0) and np.all(w[:, :, -1] == 0)  # border entries stay zero
    assert np.array_equal(PO.add_vorticity(v, 0.1, dx, 2.0)[inner], v[inner])


def test_confinement_force_of_a_shear():
    """v_z = a (y dx)^2 / 2: the curl is (a y dx, 0, 0) where both y
    neighbours are interior, |w| grows with y, so N = e_y and the force
    (N x w) dx eps = (0, 0, -w_x dx eps): velocity z changes by
    -dt a y dx dx eps."""
    Z, Y, X, dx, a, eps, dt = 9, 14, 8, 1.0 / 14, 3.0, 2.0, 0.1
    z, y, x = grid(Z, Y, X)
    v = np.stack([np.zeros_like(y), np.zeros_like(y), a * (y * dx) ** 2 / 2], axis=-1)
    w, wmag = PO.curl(v, 1.0 / dx)
    mid = (slice(1, -1), slice(2, Y - 2), slice(1, -1))
    assert np.allclose(w[mid][..., 0], (a * y * dx)[mid], rtol=1e-12)
    assert np.allclose(w[mid][..., 1:], 0, atol=1e-12)
    f = PO.confinement_force(w, wmag, 1.0 / dx, dx, eps)
    core = (slice(1, -1), slice(3, Y - 3), slice(1, -1))
    assert np.allclose(f[core][..., 2], -(a * y * dx)[core] * dx * eps, rtol=1e-12)
    assert np.allclose(f[core][..., :2], 0, atol=1e-12)
    after = PO.add_vorticity(v, dt, dx, eps)
    assert np.allclose(after[core][..., 2] - v[core][..., 2], -dt * (a * y * dx)[core] * dx * eps, rtol=1e-12)


def test_velocity_border_rules():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((6, 7, 8, 3))
    b = PO.velocity_border(v, open_x=False, open_y=True, open_z=False)
    assert np.array_equal(b[1:-1, 1:-1, 1:-1], v[1:-1, 1:-1, 1:-1])
    assert np.all(b[:, :, 0, 0] == 0) and np.all(b[:, :, -1, 0] == 0)       # setZeroX
    assert np.array_equal(b[:, 0, :, 1], v[:, 1, :, 1])                       # copyBorderY
    assert np.array_equal(b[:, -1, :, 1], v[:, -2, :, 1])
    assert np.all(b[0, :, :, 2] == 0) and np.all(b[-1, :, :, 2] == 0)       # setZeroZ
    # away from the edges, the tangential components on a face are untouched
    assert np.array_equal(b[1:-1, 1:-1, 0, 1:], v[1:-1, 1:-1, 0, 1:])


def test_smoke_column_cells():
    """48 x 64 x 48: dx = 1/64, xTotal = zTotal = 0.75, the column's axis at
    x = 0.3 / dx = 19.2, z = 0.375 / dx = 24, radius 0.075 * 0.75 / dx = 3.6
    cells, layers y = (int)(0.05 * 64) = 3 .. (int)(0.10 * 64) = 6."""
    d, h = PO.add_smoke(np.zeros((48, 64, 48)), np.zeros((48, 64, 48)))
    assert np.array_equal(d, h)
    for y in (3, 6):
        assert d[24, y, 19] == 1.0 and d[24, y, 16] == 1.0      # 0.2 and 3.2 cells from the axis
        assert d[24, y, 23] == 0.0 and d[28, y, 19] == 0.0      # 3.8 and 4.0 cells away
    assert np.all(d[:, :3] == 0) and np.all(d[:, 7:] == 0)
    assert np.array_equal(d[:, 3], d[:, 6])
