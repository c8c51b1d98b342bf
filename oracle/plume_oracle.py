"""FP64 CPU oracle of the plume timestep (stepPlume and addSmokeColumn).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  Nothing in
cfs_plume_amd/ imports it.

Fields are full grids (Z, Y, X) x fastest; velocity (Z, Y, X, 3).

set_zero_border()     VECTOR3_FIELD_3D::setZeroBorder / FIELD_3D::setZeroBorder
                      (VECTOR3_FIELD_3D.cpp:2828-2866, FIELD_3D.cpp:6408-6440)
add_buoyancy()        FLUID_3D_MIC::addBuoyancy (FLUID_3D_MIC.cpp:491-502) and the
                      step's velocity.axpy(dt, force) (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:463-469)
add_vorticity()       FLUID_3D_MIC::addVorticity (FLUID_3D_MIC.cpp:524-593): the curl
                      with one-sided differences next to the border, its magnitude,
                      N = grad|w| / |grad|w||, force = (N x w) dx eps, velocity += dt force
velocity_border()     copyBorderX/Y/Z or setZeroX/Y/Z per axis (VECTOR3_FIELD_3D.cpp:2872-3057),
                      as advectHeatAndDensityStam chooses them (:796-824)
add_smoke()           FLUID_3D_MIC::addSmokeTestCase (FLUID_3D_MIC.cpp:674-703)
plume_step()          stepPlume (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501) with the
                      codec oracle's project / unproject, closure_oracle.advect_scalar
                      and a given subspace advection.

Pins (tests/test_plume_oracle.py): the curl of a rigid rotation is 2 Omega
everywhere (central and one-sided differences are exact on linear fields);
a shear v_z = a y^2 / 2 has curl (a y, 0, 0) and confinement force
-a y dx^2 eps e_z away from the border (the sign of N x w); the border
rules leave the interior untouched; the smoke column holds the cells whose
single-precision radius lies inside 0.075 xTotal.
"""
from __future__ import annotations

import numpy as np

from oracle import closure_oracle as CO


def set_zero_border(f: np.ndarray) -> np.ndarray:
    g = np.array(f, np.float64, copy=True)
    g[0], g[-1] = 0.0, 0.0
    g[:, 0], g[:, -1] = 0.0, 0.0
    g[:, :, 0], g[:, :, -1] = 0.0, 0.0
    return g


def add_buoyancy(vel: np.ndarray, heat: np.ndarray, dt: float, beta: float) -> np.ndarray:
    v = np.array(vel, np.float64, copy=True)
    v[..., 1] += dt * (beta * heat)
    return v


def _neighbours(Z, Y, X):
    """Index arrays of the stencil neighbours of the interior cells and the
    inverse spacings (1/dx one-sided at the layer next to the border)."""
    z, y, x = np.meshgrid(np.arange(1, Z - 1), np.arange(1, Y - 1), np.arange(1, X - 1), indexing="ij")
    up = np.where(y == Y - 2, y, y + 1)
    down = np.where(y == 1, y, y - 1)
    out = np.where(z == Z - 2, z, z + 1)
    inn = np.where(z == 1, z, z - 1)
    right = np.where(x == X - 2, x, x + 1)
    left = np.where(x == 1, x, x - 1)
    one_y = (up == y) | (down == y)
    one_z = (out == z) | (inn == z)
    one_x = (right == x) | (left == x)
    return z, y, x, up, down, out, inn, right, left, one_y, one_z, one_x


def curl(vel: np.ndarray, inv_dx: float):
    """The vorticity w and |w| of addVorticity's first loop (border entries 0)."""
    Z, Y, X, _ = vel.shape
    z, y, x, up, down, out, inn, right, left, one_y, one_z, one_x = _neighbours(Z, Y, X)
    grid = 0.5 * inv_dx
    dy = np.where(one_y, inv_dx, grid)
    dz = np.where(one_z, inv_dx, grid)
    dxs = np.where(one_x, inv_dx, grid)
    v = vel
    w = np.zeros_like(vel, dtype=np.float64)
    w[z, y, x, 0] = (v[z, up, x, 2] - v[z, down, x, 2]) * dy + (-v[out, y, x, 1] + v[inn, y, x, 1]) * dz
    w[z, y, x, 1] = (v[out, y, x, 0] - v[inn, y, x, 0]) * dz + (-v[z, y, right, 2] + v[z, y, left, 2]) * dxs
    w[z, y, x, 2] = (v[z, y, right, 1] - v[z, y, left, 1]) * dxs + (-v[z, up, x, 0] + v[z, down, x, 0]) * dy
    return w, np.linalg.norm(w, axis=-1)


def confinement_force(w: np.ndarray, wmag: np.ndarray, inv_dx: float, dx: float, eps: float) -> np.ndarray:
    """addVorticity's second loop: the force at interior cells."""
    Z, Y, X, _ = w.shape
    z, y, x, up, down, out, inn, right, left, one_y, one_z, one_x = _neighbours(Z, Y, X)
    grid = 0.5 * inv_dx
    n0 = (wmag[z, y, right] - wmag[z, y, left]) * np.where(one_x, inv_dx, grid)
    n1 = (wmag[z, up, x] - wmag[z, down, x]) * np.where(one_y, inv_dx, grid)
    n2 = (wmag[out, y, x] - wmag[inn, y, x]) * np.where(one_z, inv_dx, grid)
    m = np.sqrt(n0 * n0 + n1 * n1 + n2 * n2)
    ok = m > 0.0
    inv = np.where(ok, 1.0 / np.where(ok, m, 1.0), 0.0)
    n0, n1, n2 = n0 * inv, n1 * inv, n2 * inv
    w0, w1, w2 = w[z, y, x, 0], w[z, y, x, 1], w[z, y, x, 2]
    f = np.zeros_like(w)
    f[z, y, x, 0] = np.where(ok, (n1 * w2 - n2 * w1) * dx * eps, 0.0)
    f[z, y, x, 1] = np.where(ok, -(n0 * w2 - n2 * w0) * dx * eps, 0.0)
    f[z, y, x, 2] = np.where(ok, (n0 * w1 - n1 * w0) * dx * eps, 0.0)
    return f


def add_vorticity(vel: np.ndarray, dt: float, dx: float, eps: float) -> np.ndarray:
    if eps <= 0.0:
        return np.array(vel, np.float64, copy=True)
    w, wmag = curl(vel, 1.0 / dx)
    return vel + dt * confinement_force(w, wmag, 1.0 / dx, dx, eps)


def velocity_border(vel: np.ndarray, open_x: bool, open_y: bool, open_z: bool) -> np.ndarray:
    v = np.array(vel, np.float64, copy=True)
    if open_x:
        v[:, :, 0, 0], v[:, :, -1, 0] = v[:, :, 1, 0], v[:, :, -2, 0]
    else:
        v[:, :, 0, 0], v[:, :, -1, 0] = 0.0, 0.0
    if open_y:
        v[:, 0, :, 1], v[:, -1, :, 1] = v[:, 1, :, 1], v[:, -2, :, 1]
    else:
        v[:, 0, :, 1], v[:, -1, :, 1] = 0.0, 0.0
    if open_z:
        v[0, :, :, 2], v[-1, :, :, 2] = v[1, :, :, 2], v[-2, :, :, 2]
    else:
        v[0, :, :, 2], v[-1, :, :, 2] = 0.0, 0.0
    return v


def add_smoke(density: np.ndarray, heat: np.ndarray):
    """addSmokeTestCase on density and heat (Z, Y, X)."""
    Z, Y, X = density.shape
    maxres = max(X, Y, Z)
    dx = 1.0 / float(np.float32(maxres))
    xtot, ztot = dx * X, dx * Z
    d, h = np.array(density, np.float64), np.array(heat, np.float64)
    for y in range(int(0.05 * Y), int(0.10 * Y) + 1):
        for z in range(Z):
            for x in range(X):
                xl = x * dx - xtot * float(np.float32(0.4))
                zl = z * dx - ztot * float(np.float32(0.5))
                radius = float(np.sqrt(np.float32(xl * xl + zl * zl)))
                if radius < float(np.float32(0.075)) * xtot:
                    d[z, y, x] = 1.0
                    h[z, y, x] = 1.0
    return d, h


def plume_step(pre, fin, D, P, I, params, vel, density, heat, advect):
    """stepPlume on full-grid FP64 fields; pre/fin are oracle.Basis objects,
    advect(q) -> q' the subspace advection.  Returns (vel, density, heat, q)."""
    Z, Y, X = density.shape
    dt, dx = params["dt"], params["dx"]
    v = set_zero_border(vel)
    v = add_buoyancy(v, heat, dt, params["buoyancy"])
    v = add_vorticity(v, dt, dx, params["vorticity_eps"])
    q = pre.project(v.reshape(-1))
    v = velocity_border(v, params["open_x"], params["open_y"], params["open_z"])
    full = (X, Y, Z)
    d = CO.advect_scalar(v.reshape(-1), density.reshape(-1), full, dt / dx).reshape(Z, Y, X)
    h = CO.advect_scalar(v.reshape(-1), heat.reshape(-1), full, dt / dx).reshape(Z, Y, X)
    d, h = set_zero_border(d), set_zero_border(h)
    q = advect(q)
    q = D @ q
    q = P @ q + I @ q
    v = fin.unproject(q, v.reshape(-1)).reshape(Z, Y, X, 3)
    return v, d, h, q
