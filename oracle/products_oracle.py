"""FP64 CPU oracle of the basis builder's sparse operators and products.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  Nothing in
cfs_plume_amd/ imports it.

Sparse matrices on the peeled grid (scipy CSR), built with the reference's
loops (src/integrators/SUBSPACE_FLUID_3D_EIGEN.cpp):

poisson_flat()       buildFlatA (:2166-2207)
vel_to_div()         computeVelocityToDivergence (:1989-2100)
pressure_to_vel()    computePressureToVelocity (:627-695)
damping_flat()       buildPeeledDampingMatrixFlat (:2107-2160)
products()           buildOutOfCoreMatrices (:1355-1489) from dense bases

Pins (tests/test_products_oracle.py): each operator agrees with the
full-rank solver's own restatement of the same stage (oracle/fluid_oracle.py:
the Poisson matrix of buildA, the divergence and gradient of project() with
the Neumann/wall borders, the full damping matrix), written from different
reference functions.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import fluid_oracle as FO


def _peeled(full):
    X, Y, Z = full
    return X - 2, Y - 2, Z - 2


def poisson_flat(full, bc=FO.WALLS):
    X, Y, Z = full
    Xp, Yp, Zp = _peeled(full)
    slab, sp_ = X * Y, Xp * Yp
    skip = FO.obstacles(full, bc)
    rows, cols, vals = [], [], []
    diag = {}
    for z in range(1, Z - 1):
        for y in range(1, Y - 1):
            for x in range(1, X - 1):
                i = x + y * X + z * slab
                if skip[i]:
                    continue
                p = (x - 1) + (y - 1) * Xp + (z - 1) * sp_
                diag[p] = sum(1.0 for j in (i + 1, i - 1, i + X, i - X, i + slab, i - slab) if not skip[j])
                for cond, j, pj in ((x != 1, i - 1, p - 1), (x != X - 2, i + 1, p + 1), (y != 1, i - X, p - Xp),
                                    (y != Y - 2, i + X, p + Xp), (z != 1, i - slab, p - sp_),
                                    (z != Z - 2, i + slab, p + sp_)):
                    if cond and not skip[j]:
                        rows.append(p)
                        cols.append(pj)
                        vals.append(-1.0)
    for p, v in diag.items():
        rows.append(p)
        cols.append(p)
        vals.append(v)
    n = Xp * Yp * Zp
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def vel_to_div(full, dx, bc=FO.WALLS):
    """bc = front, back, left, right, top, bottom (1 wall)."""
    Xp, Yp, Zp = _peeled(full)
    sp_ = Xp * Yp
    coeff = -dx * 0.5
    front, back, left_bc, right_bc, top, bottom = bc
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r)
        cols.append(c)
        vals.append(v)

    for z in range(Zp):
        for y in range(Yp):
            for x in range(Xp):
                i = x + y * Xp + z * sp_
                right, left, up, down, far, near = i + 1, i - 1, i + Xp, i - Xp, i + sp_, i - sp_
                if x != Xp - 1:
                    add(i, 3 * right, coeff)
                elif right_bc != 0:
                    add(i, 3 * i, -coeff)
                else:
                    add(i, 3 * left, coeff)
                if x != 0:
                    add(i, 3 * left, -coeff)
                elif left_bc != 0:
                    add(i, 3 * i, coeff)
                else:
                    add(i, 3 * right, -coeff)
                if y != Yp - 1:
                    add(i, 3 * up + 1, coeff)
                elif top != 0:
                    add(i, 3 * i + 1, -coeff)
                else:
                    add(i, 3 * down + 1, coeff)
                if y != 0:
                    add(i, 3 * down + 1, -coeff)
                elif bottom != 0:
                    add(i, 3 * i + 1, coeff)
                else:
                    add(i, 3 * up + 1, -coeff)
                if z != Zp - 1:
                    add(i, 3 * far + 2, coeff)
                elif back != 0:
                    add(i, 3 * i + 2, -coeff)
                else:
                    add(i, 3 * near + 2, coeff)
                if z != 0:
                    add(i, 3 * near + 2, -coeff)
                elif front != 0:
                    add(i, 3 * i + 2, coeff)
                else:
                    add(i, 3 * far + 2, -coeff)
    n = Xp * Yp * Zp
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, 3 * n))  # duplicates are summed


def pressure_to_vel(full, dx):
    Xp, Yp, Zp = _peeled(full)
    sp_ = Xp * Yp
    inv, half = 1.0 / dx, 0.5 / dx
    rows, cols, vals = [], [], []
    for z in range(Zp):
        for y in range(Yp):
            for x in range(Xp):
                i = x + y * Xp + z * sp_
                for c, (pos, lim, s) in enumerate(((x, Xp, 1), (y, Yp, Xp), (z, Zp, sp_))):
                    if pos == 0:
                        ent = ((i + s, -inv), (i, inv))
                    elif pos == lim - 1:
                        ent = ((i, -inv), (i - s, inv))
                    else:
                        ent = ((i + s, -half), (i - s, half))
                    for col, v in ent:
                        rows.append(3 * i + c)
                        cols.append(col)
                        vals.append(v)
    n = Xp * Yp * Zp
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * n, n))


def damping_flat(full, w=0.9):
    Xp, Yp, Zp = _peeled(full)
    sp_ = Xp * Yp
    sixth, keep = (1.0 / 6.0) * w, 1.0 - w
    rows, cols, vals = [], [], []
    for z in range(Zp):
        for y in range(Yp):
            for x in range(Xp):
                i = x + y * Xp + z * sp_
                nb = [(x != Xp - 1, i + 1), (x != 0, i - 1), (y != Yp - 1, i + Xp), (y != 0, i - Xp),
                      (z != Zp - 1, i + sp_), (z != 0, i - sp_)]
                for c in range(3):
                    rows.append(3 * i + c)
                    cols.append(3 * i + c)
                    vals.append(keep)
                    for ok, j in nb:
                        if ok:
                            rows.append(3 * i + c)
                            cols.append(3 * j + c)
                            vals.append(sixth)
    n = 3 * Xp * Yp * Zp
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def products(full, dx, U, bc=FO.WALLS, w=0.9):
    """U: dict of dense bases (rows x r) 'pressure', 'preproject', 'prediffuse', 'final'."""
    Up, Uproj, Udiff, Ufin = U["pressure"], U["preproject"], U["prediffuse"], U["final"]
    A = Up.T @ (poisson_flat(full, bc) @ Up)
    vtod = Up.T @ (vel_to_div(full, dx, bc) @ Uproj)
    damping = Uproj.T @ (damping_flat(full, w) @ Udiff)
    ptof = Ufin.T @ Uproj
    ptov = Ufin.T @ (pressure_to_vel(full, dx) @ Up)
    return {"projected.A": A, "projected.vtod": vtod, "damping.peeled": damping, "projected.ptof": ptof,
            "projected.ptov": ptov, "inverseProduct": ptov @ np.linalg.inv(A) @ vtod}
