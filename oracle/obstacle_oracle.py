"""FP64 CPU oracle of the moving-obstacle steps (a spinning, sliding box).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  Nothing in
cfs_plume_amd/ imports it.

Box                  BOX (src/geometry/BOX.h, BOX.cpp:21-131): a box of given
                     centre and edge lengths in the unit cube, spinning about
                     the z axis with `period` and sliding along x as
                     0.25 sin(phi); inside() rotates the point back by the
                     rotation R = rotation(z, -theta) (MATRIX3.h:96-117);
                     velocity_at() is update_r + update_linearVelocity
                     (omega x r + the sliding velocity); advance() is
                     translate_center, spin, update_time (the per-step order of
                     SUBSPACE_FLUID_3D_EIGEN.cpp:1715-1720)
cell_centers()       FLUID_3D_MIC::cellCenter (FLUID_3D_MIC.cpp:1704-1733)
stomp()              setPeeledSparseMovingIOP + setVelocityNeumann and the
                     product neumannIOP [v; 1] (FLUID_3D_MIC.cpp:1645-1702,
                     :279-282): peeled components of cells inside the box take
                     the box's velocity there
complement()         setPeeledSparseMovingIOPComplement
                     (SUBSPACE_FLUID_3D_EIGEN.cpp:713-758): the diagonal of the
                     complement matrix C and the Neumann vector
reduced_iop()        buildObstacleMatrices' per-step pair
                     (SUBSPACE_FLUID_3D_EIGEN.cpp:1699-1711):
                     R = U_iop^T U_proj - U_iop^T C U_proj, n = U_iop^T nvec
step_moving_obstacle()  FLUID_3D_MIC::stepWithMovingObstacle (:236-310) with the
                     snapshots of appendStreams, IOP stream included (:1376-1480)
reduced_step()       SUBSPACE_FLUID_3D_COMPRESSED_EIGEN::stepMovingObstacle
                     (:508-566): project, scalar advection, subspace advection,
                     D, q = R q + n (reducedSetMovingBox :779-790), P + I,
                     unproject

Pins (tests/test_obstacle_oracle.py): the rotation matrix is orthonormal with
det 1 and turns e_x into (cos theta, -sin theta, 0); at theta = pi/2 the box's
x and y extents swap; velocity_at() is a rigid rotation about the centre's
z axis plus the slide (tangential, |v - v_slide| = omega |r_xy|); the time
schedule reproduces 0.25 sin(2 pi t / T) and the wrap of theta; the reduced
pair (R, n) equals U_iop^T stomp(U_proj q) for any q -- the reference's
subspace formula against its full-space stomp, two different functions;
the stomped cells carry the box velocity and the rest are untouched.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import closure_oracle as CO
from oracle import fluid_oracle as FO
from oracle import plume_oracle as PO

# the box of projects/buildProducts/buildProducts.cpp:11-18
DEFAULT_CENTER = (0.5, 0.5, 0.5)
DEFAULT_LENGTHS = (0.4, 0.05, 0.15)
DEFAULT_PERIOD = 4.0
DEFAULT_TRANSLATION_PERIOD = 10.0


def rotation(axis, theta):
    """MATRIX3::rotation (MATRIX3.h:96-117): rotation by theta about axis."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(theta), math.sin(theta)
    v = 1.0 - c
    x, y, z = a
    return np.array([[x * x * v + c, x * y * v - z * s, x * z * v + y * s],
                     [y * x * v + z * s, y * y * v + c, y * z * v - x * s],
                     [z * x * v - y * s, z * y * v + x * s, z * z * v + c]])


class Box:
    def __init__(self, center=DEFAULT_CENTER, lengths=DEFAULT_LENGTHS, period=DEFAULT_PERIOD,
                 translation_period=DEFAULT_TRANSLATION_PERIOD, dt=0.1):
        self.center = np.array(center, np.float64)
        self.original_center = self.center.copy()
        self.half = 0.5 * np.array(lengths, np.float64)
        self.period = float(period)
        self.translation_period = float(translation_period)
        self.omega = np.array([0.0, 0.0, 2.0 * math.pi / self.period])
        self.axis = np.array([0.0, 0.0, 1.0])
        self.theta = 0.0
        self.phi = 0.0
        self.time = 0.0
        self.dt = float(dt)
        self.R = np.eye(3)

    def inside(self, points):
        """points (N, 3) -> bool (N,); R (p - c) summed left to right per row."""
        u = np.asarray(points, np.float64) - self.center
        R = self.R
        t = [R[i, 0] * u[:, 0] + R[i, 1] * u[:, 1] + R[i, 2] * u[:, 2] for i in range(3)]
        return (np.abs(t[0]) <= self.half[0]) & (np.abs(t[1]) <= self.half[1]) & (np.abs(t[2]) <= self.half[2])

    def slide_velocity(self):
        return np.array([0.25 * 2.0 * math.pi / self.translation_period * math.cos(self.phi), 0.0, 0.0])

    def velocity_at(self, points):
        u = np.asarray(points, np.float64) - self.center
        r = u - np.outer(u @ self.axis / (self.axis @ self.axis), self.axis)
        return np.cross(self.omega, r) + self.slide_velocity()

    def advance(self):
        # translate_center
        self.phi += 2.0 * math.pi * self.dt / self.translation_period
        if self.phi >= 2.0 * math.pi:
            self.phi = 0.0
        self.center[0] = self.original_center[0] + 0.25 * math.sin(self.phi)
        # spin
        self.theta += 2.0 * math.pi * self.dt / self.period
        if self.theta >= 2.0 * math.pi:
            self.theta = 0.0
        self.R = rotation(self.axis, -self.theta)
        # update_time
        self.time += self.dt


def cell_centers(full):
    """Centres of the peeled cells (x fastest), full-grid indices 1..res-2."""
    X, Y, Z = full
    z, y, x = np.meshgrid(np.arange(1, Z - 1), np.arange(1, Y - 1), np.arange(1, X - 1), indexing="ij")
    dx, dy, dz = 1.0 / X, 1.0 / Y, 1.0 / Z
    return np.stack([x.reshape(-1) * dx + dx * 0.5, y.reshape(-1) * dy + dy * 0.5, z.reshape(-1) * dz + dz * 0.5],
                    axis=1)


def complement(full, box):
    """(mask over peeled cells, Neumann vector over peeled components)."""
    pts = cell_centers(full)
    inside = box.inside(pts)
    nvec = np.zeros((len(pts), 3))
    nvec[inside] = box.velocity_at(pts[inside])
    return inside, nvec.reshape(-1)


def stomp_peeled(full, box, vp):
    inside, nvec = complement(full, box)
    out = np.array(vp, np.float64, copy=True).reshape(-1, 3)
    out[inside] = nvec.reshape(-1, 3)[inside]
    return out.reshape(-1)


def stomp(full, box, vel):
    """vel: flat full-grid interleaved; the interior set from the stomped peeled field."""
    X, Y, Z = full
    v = np.array(vel, np.float64, copy=True).reshape(Z, Y, X, 3)
    vp = stomp_peeled(full, box, FO.peel(full, v, 3))
    v[1:-1, 1:-1, 1:-1] = vp.reshape(Z - 2, Y - 2, X - 2, 3)
    return v.reshape(-1)


def reduced_iop(full, box, U_iop, U_proj):
    """(R, n): R = U_iop^T U_proj - U_iop^T C U_proj, n = U_iop^T nvec; U (3 Np x r)."""
    inside, nvec = complement(full, box)
    c = np.repeat(inside, 3).astype(np.float64)
    R = U_iop.T @ U_proj - U_iop.T @ (c[:, None] * U_proj)
    return R, U_iop.T @ nvec


def step_moving_obstacle(solver, vel, dens, heat, box):
    """FLUID_3D_MIC::stepWithMovingObstacle on an oracle fluid_oracle.Solver:
    no buoyancy; vorticity, advectStam, damping, the box stomp, project.
    Returns (vel, dens, heat, snapshots, (iterations, max residual))."""
    full, P = solver.full, solver.prm
    X, Y, Z = full
    shp = (Z, Y, X)
    dt, dx = P["dt"], P["dx"]
    v = PO.set_zero_border(np.asarray(vel).reshape(*shp, 3))
    d = PO.set_zero_border(np.asarray(dens).reshape(shp)).reshape(-1)
    h = np.asarray(heat, np.float64).reshape(-1)
    if P["vorticity_eps"] != 0.0:
        v = PO.add_vorticity(v, dt, dx, P["vorticity_eps"])
    v = v.reshape(-1)
    snaps = {}
    v = FO.border(full, v, solver.open, 1)
    snaps["preadvect"] = FO.peel(full, v, 3)
    dt0 = dt / dx
    d = CO.advect_scalar(v, d, full, dt0).reshape(-1)
    h = CO.advect_scalar(v, h, full, dt0).reshape(-1)
    v = CO.advect(v, full, dt0).reshape(-1)
    v = PO.set_zero_border(v.reshape(*shp, 3)).reshape(-1)
    v = FO.border(full, v, solver.open, 1)
    d = PO.set_zero_border(d.reshape(shp)).reshape(-1)
    h = PO.set_zero_border(h.reshape(shp)).reshape(-1)
    snaps["prediffuse"] = FO.peel(full, v, 3)
    v = FO.damping(full, v, P["damping_w"])
    snaps["preproject"] = FO.peel(full, v, 3)
    v = stomp(full, box, v)
    snaps["iop"] = FO.peel(full, v, 3)
    v = FO.border(full, v, solver.open, 2)
    div = FO.divergence(full, v, solver.skip, dx)
    p, its, mx = FO.solve_poisson(full, solver.A, solver.pc, solver.skip, div, P["solver_eps"], P["iterations"],
                                  solver.lv)
    v = FO.gradient_subtract(full, v, p, dx)
    snaps["final_velocity"] = FO.peel(full, v, 3)
    snaps["pressure"] = FO.peel(full, p, 1)
    snaps["divergence"] = FO.peel(full, div, 1)
    return v, d, h, snaps, (its, mx)


def reduced_step(project, unproject, D, R, n, P, I, params, vel, density, heat, advect):
    """stepMovingObstacle on full-grid FP64 fields (Z, Y, X[, 3]).
    project(v_flat) -> q in U_preadvect coordinates; unproject(q, v_flat) ->
    v_flat with the interior from U_final; advect(q) -> q'.
    Returns (vel, density, heat, q)."""
    Z, Y, X = density.shape
    dt, dx = params["dt"], params["dx"]
    v = PO.set_zero_border(vel)
    v = PO.add_vorticity(v, dt, dx, params["vorticity_eps"])
    q = project(v.reshape(-1))
    v = PO.velocity_border(v, params["open_x"], params["open_y"], params["open_z"])
    full = (X, Y, Z)
    d = CO.advect_scalar(v.reshape(-1), density.reshape(-1), full, dt / dx).reshape(Z, Y, X)
    h = CO.advect_scalar(v.reshape(-1), heat.reshape(-1), full, dt / dx).reshape(Z, Y, X)
    d, h = PO.set_zero_border(d), PO.set_zero_border(h)
    q = advect(q)
    q = D @ q
    q = R @ q + n
    q = P @ q + I @ q
    v = unproject(q, v.reshape(-1)).reshape(Z, Y, X, 3)
    return v, d, h, q
