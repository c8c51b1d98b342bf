"""FP64 CPU oracle of the full-rank Stam plume (FLUID_3D_MIC) for small grids.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  Nothing in
cfs_plume_amd/ imports it.

Fields are flat full-grid arrays, index x + X (y + Y z); velocity flat
interleaved xyz.  Follows src/integrators/FLUID_3D_MIC.cpp:

obstacles()        the domain walls of the constructor (:100-160)
build_a()          buildA (:772-801), skip = obstacles, heat = false
build_precon()     buildMICPreconditioner (:825-862), tau 0.97, sigma 0.25
apply_mic()        applyMICPreconditioner (:864-899): forward L q = r, backward L^T z = q
mult_a()           multA (:803-823)
solve_poisson()    solvePoisson (:918-1017): MIC-PCG, stop at max(r) <= eps (signed max,
                   FIELD_3D::max FIELD_3D.cpp:1529-1538), the iteration cap, or a
                   relative change of max(r) below 1e-7
damping()          buildPeeledDampingMatrixFull (:1251-1307) times the peeled velocity
project()          project (:314-402)
step()             step (:179-233) with the snapshots of appendStreams (:1329-1480)

Pins (tests/test_fluid_oracle.py): the MIC sweeps equal the triangular
solves of L = (diag(1/pc) + strictly-lower A) diag(pc) built as a sparse
matrix and solved by scipy; the CG pressure solves A p = b to the stated
tolerance against scipy's direct solve; the projected velocity's divergence
(the reference's own stencil) is at the solver tolerance; interior rows of
the damping matrix sum to one (a constant field stays constant away from
the border); a fluid at rest without smoke stays at rest.
"""
from __future__ import annotations

import numpy as np

from oracle import closure_oracle as CO
from oracle import plume_oracle as PO

WALLS = (1, 1, 1, 1, 0, 0)  # front, back (z), left, right (x), top, bottom (y)


def default_params(full, vorticity=0.0):
    X, Y, Z = full
    maxres = max(full)
    return {"dt": 0.1, "dx": 1.0 / maxres, "buoyancy": float(np.float32(0.1)), "vorticity_eps": vorticity,
            "damping_w": 0.9, "solver_eps": 1e-10, "iterations": 1000, "boundaries": WALLS}


def obstacles(full, bc=WALLS) -> np.ndarray:
    X, Y, Z = full
    o = np.zeros((Z, Y, X), bool)
    if bc[0] == 1:
        o[0] = True
    if bc[1] == 1:
        o[-1] = True
    if bc[5] == 1:
        o[:, 0] = True
    if bc[4] == 1:
        o[:, -1] = True
    if bc[2] == 1:
        o[:, :, 0] = True
    if bc[3] == 1:
        o[:, :, -1] = True
    return o.reshape(-1)


def _interior(full):
    X, Y, Z = full
    z, y, x = np.meshgrid(np.arange(1, Z - 1), np.arange(1, Y - 1), np.arange(1, X - 1), indexing="ij")
    return (x + X * (y + Y * z)).reshape(-1)


def build_a(full, skip):
    X, Y, Z = full
    n, slab = X * Y * Z, X * Y
    Ad, Api, Apj, Apk = (np.zeros(n) for _ in range(4))
    for i in _interior(full):
        if skip[i]:
            continue
        Ad[i] = sum(1.0 for j in (i + 1, i - 1, i + X, i - X, i + slab, i - slab) if not skip[j])
        Api[i - 1] = 0.0 if skip[i - 1] else -1.0
        Api[i] = 0.0 if skip[i + 1] else -1.0
        Apj[i - X] = 0.0 if skip[i - X] else -1.0
        Apj[i] = 0.0 if skip[i + X] else -1.0
        Apk[i - slab] = 0.0 if skip[i - slab] else -1.0
        Apk[i] = 0.0 if skip[i + slab] else -1.0
    return Ad, Api, Apj, Apk


def build_precon(full, A, tau=0.97, sig=0.25):
    X, Y, Z = full
    slab = X * Y
    Ad, Api, Apj, Apk = A
    pc = np.zeros_like(Ad)
    for i in _interior(full):
        xm, ym, zm = i - 1, i - X, i - slab
        pi2, pj2, pk2 = pc[xm] ** 2, pc[ym] ** 2, pc[zm] ** 2
        e0 = Api[xm] * Api[xm] * pi2 + Apj[ym] * Apj[ym] * pj2 + Apk[zm] * Apk[zm] * pk2
        e1 = tau * (Api[xm] * (Apj[xm] + Apk[xm]) * pi2 + Apj[ym] * (Api[ym] + Apk[ym]) * pj2
                    + Apk[zm] * (Api[zm] + Apj[zm]) * pk2)
        e = Ad[i] - e0 - e1
        if e < sig * Ad[i]:
            e = Ad[i]
        pc[i] = 0.0 if abs(Ad[i]) < 0.000001 else 1.0 / np.sqrt(e)
    return pc


def levels(full):
    """Interior cell indices grouped by x + y + z (the sweeps' dependency levels)."""
    X, Y, Z = full
    idx = _interior(full)
    x, y, z = idx % X, (idx // X) % Y, idx // (X * Y)
    s = x + y + z
    return [idx[s == k] for k in range(3, s.max() + 1)]


def apply_mic(full, A, pc, r, lv=None):
    X, Y, Z = full
    slab = X * Y
    _, Api, Apj, Apk = A
    lv = lv if lv is not None else levels(full)
    q = np.zeros_like(r)
    zf = np.zeros_like(r)
    for i in lv:  # cells of one level depend only on earlier levels
        xm, ym, zm = i - 1, i - X, i - slab
        t = r[i] - Api[xm] * pc[xm] * q[xm] - Apj[ym] * pc[ym] * q[ym] - Apk[zm] * pc[zm] * q[zm]
        q[i] = t * pc[i]
    for i in reversed(lv):
        p = pc[i]
        t = q[i] - Api[i] * p * zf[i + 1] - Apj[i] * p * zf[i + X] - Apk[i] * p * zf[i + slab]
        zf[i] = t * p
    return zf


def mult_a(full, A, xf):
    X, Y, Z = full
    slab = X * Y
    Ad, Api, Apj, Apk = A
    i = _interior(full)
    b = np.zeros_like(xf)
    b[i] = (Ad[i] * xf[i] + xf[i - 1] * Api[i - 1] + xf[i + 1] * Api[i] + xf[i - X] * Apj[i - X]
            + xf[i + X] * Apj[i] + xf[i - slab] * Apk[i - slab] + xf[i + slab] * Apk[i])
    return b


def solve_poisson(full, A, pc, skip, b, eps=1e-10, iterations=1000, lv=None):
    """solvePoisson from x = 0; returns (x, iterations, max residual)."""
    x = np.zeros_like(b)
    inner = _interior(full)
    r = np.zeros_like(b)
    r[inner] = b[inner] - mult_a(full, A, x)[inner]
    r[skip] = 0.0
    z = apply_mic(full, A, pc, r, lv)
    d = z.copy()
    delta_new = r @ z
    max_r = 2.0 * eps
    i = 0
    while i < iterations and max_r > eps:
        z = mult_a(full, A, d)
        alpha = d @ z
        if abs(alpha) > 0.0:
            alpha = delta_new / alpha
        else:
            break
        x += alpha * d
        r += -alpha * z
        z = apply_mic(full, A, pc, r, lv)
        old = max_r
        max_r = r.max()
        delta_old = delta_new
        delta_new = r @ z
        beta = delta_new / delta_old
        d = d * beta + z
        d[skip] = 0.0
        i += 1
        if abs((max_r - old) / old) < 1e-7:
            break
    return x, i, max_r


def damping(full, vel, w=0.9):
    """(1 - w) v + (w / 6) sum of the neighbours inside the peeled domain, at interior cells."""
    X, Y, Z = full
    v = np.asarray(vel, np.float64).reshape(Z, Y, X, 3)
    p = v[1:-1, 1:-1, 1:-1]
    sixth, keep = (1.0 / 6.0) * w, 1.0 - w
    out = v.copy()
    acc = np.zeros_like(p)
    acc[1:] += p[:-1] * sixth
    acc[:, 1:] += p[:, :-1] * sixth
    acc[:, :, 1:] += p[:, :, :-1] * sixth
    acc += p * keep
    acc[:, :, :-1] += p[:, :, 1:] * sixth
    acc[:, :-1] += p[:, 1:] * sixth
    acc[:-1] += p[1:] * sixth
    out[1:-1, 1:-1, 1:-1] = acc
    return out.reshape(-1)


def border(full, vel, open_xyz, reach):
    """setZero/copyBorder (reach 1) or setNeumann (reach 2) per axis."""
    X, Y, Z = full
    v = np.array(vel, np.float64).reshape(Z, Y, X, 3)
    ox, oy, oz = open_xyz
    if ox:
        v[:, :, 0, 0], v[:, :, -1, 0] = v[:, :, reach, 0], v[:, :, -1 - reach, 0]
    else:
        v[:, :, 0, 0], v[:, :, -1, 0] = 0.0, 0.0
    if oy:
        v[:, 0, :, 1], v[:, -1, :, 1] = v[:, reach, :, 1], v[:, -1 - reach, :, 1]
    else:
        v[:, 0, :, 1], v[:, -1, :, 1] = 0.0, 0.0
    if oz:
        v[0, :, :, 2], v[-1, :, :, 2] = v[reach, :, :, 2], v[-1 - reach, :, :, 2]
    else:
        v[0, :, :, 2], v[-1, :, :, 2] = 0.0, 0.0
    return v.reshape(-1)


def divergence(full, vel, skip, dx):
    X, Y, Z = full
    slab = X * Y
    v = np.asarray(vel, np.float64)
    i = _interior(full)
    xr, xl = v[3 * (i + 1)], v[3 * (i - 1)]
    yu, yd = v[3 * (i + X) + 1], v[3 * (i - X) + 1]
    zt, zb = v[3 * (i + slab) + 2], v[3 * (i - slab) + 2]
    xr = np.where(skip[i + 1], -v[3 * i], xr)
    xl = np.where(skip[i - 1], -v[3 * i], xl)
    yu = np.where(skip[i + X], -v[3 * i + 1], yu)
    yd = np.where(skip[i - X], -v[3 * i + 1], yd)
    zt = np.where(skip[i + slab], -v[3 * i + 2], zt)
    zb = np.where(skip[i - slab], -v[3 * i + 2], zb)
    div = np.zeros(X * Y * Z)
    div[i] = -dx * 0.5 * (xr - xl + yu - yd + zt - zb)
    return div


def gradient_subtract(full, vel, p, dx):
    X, Y, Z = full
    slab = X * Y
    v = np.array(vel, np.float64)
    i = _interior(full)
    x, y, z = i % X, (i // X) % Y, i // slab
    inv = 1.0 / dx
    for c, (s, coord, lim) in enumerate(((1, x, X), (X, y, Y), (slab, z, Z))):
        g = np.where(coord == 1, (p[i + s] - p[i]) * inv,
                     np.where(coord == lim - 2, (p[i] - p[i - s]) * inv, 0.5 * (p[i + s] - p[i - s]) * inv))
        v[3 * i + c] -= g
    return v


def peel(full, f, comps):
    X, Y, Z = full
    return np.asarray(f).reshape(Z, Y, X, comps)[1:-1, 1:-1, 1:-1].reshape(-1)


class Solver:
    """The matrices of one grid, built once."""

    def __init__(self, full, params=None):
        self.full = tuple(full)
        self.prm = params or default_params(full)
        self.skip = obstacles(full, self.prm["boundaries"])
        self.A = build_a(full, self.skip)
        self.pc = build_precon(full, self.A)
        self.lv = levels(full)
        bc = self.prm["boundaries"]
        self.open = (bc[2] == 0, bc[4] == 0, bc[0] == 0)

    def step(self, vel, dens, heat):
        """One FLUID_3D_MIC::step; returns (vel, dens, heat, snapshots, (iterations, max residual))."""
        full, P = self.full, self.prm
        X, Y, Z = full
        shp = (Z, Y, X)
        dt, dx = P["dt"], P["dx"]
        v = PO.set_zero_border(np.asarray(vel).reshape(*shp, 3)).reshape(-1)
        d = PO.set_zero_border(np.asarray(dens).reshape(shp)).reshape(-1)
        h = np.asarray(heat, np.float64).reshape(-1)
        v = PO.add_buoyancy(v.reshape(*shp, 3), h.reshape(shp), dt, P["buoyancy"])
        if P["vorticity_eps"] != 0.0:
            v = PO.add_vorticity(v, dt, dx, P["vorticity_eps"])
        v = v.reshape(-1)
        snaps = {}
        # advectStam
        v = border(full, v, self.open, 1)
        snaps["preadvect"] = peel(full, v, 3)
        dt0 = dt / dx
        d = CO.advect_scalar(v, d, full, dt0).reshape(-1)
        h = CO.advect_scalar(v, h, full, dt0).reshape(-1)
        v = CO.advect(v, full, dt0).reshape(-1)
        v = PO.set_zero_border(v.reshape(*shp, 3)).reshape(-1)
        v = border(full, v, self.open, 1)
        d = PO.set_zero_border(d.reshape(shp)).reshape(-1)
        h = PO.set_zero_border(h.reshape(shp)).reshape(-1)
        # diffusion
        snaps["prediffuse"] = peel(full, v, 3)
        v = damping(full, v, P["damping_w"])
        snaps["preproject"] = peel(full, v, 3)
        # project
        v = border(full, v, self.open, 2)
        div = divergence(full, v, self.skip, dx)
        p, its, mx = solve_poisson(full, self.A, self.pc, self.skip, div, P["solver_eps"], P["iterations"], self.lv)
        v = gradient_subtract(full, v, p, dx)
        snaps["final_velocity"] = peel(full, v, 3)
        snaps["pressure"] = peel(full, p, 1)
        snaps["divergence"] = peel(full, div, 1)
        return v, d, h, snaps, (its, mx)

    def add_smoke(self, dens, heat):
        X, Y, Z = self.full
        d, h = PO.add_smoke(np.asarray(dens).reshape(Z, Y, X), np.asarray(heat).reshape(Z, Y, X))
        return d.reshape(-1), h.reshape(-1)
