"""FP64 CPU oracle of the cubature reduced advection and its training columns.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  Nothing in
cfs_plume_amd/ imports it.

after_matrices():    readAdvectionCubature (reference
                     SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1262-1271): After_i =
                     w_i * U_after(x_i)^T, the r x 3 rows of the after-basis at
                     key cell x_i scaled by its weight.
cubature_advect():   reducedAdvectCompressionFriendly (:1781-1854) with
                     accumAdvectRequests (:1859-1993): the key cell's velocity
                     U_pre(x) q, the backtrace clamped to [1.5, res - 2.5], the
                     8 trilinear corners queued in a multimap keyed by block
                     number (ComputeBlockNumber, COMPRESSION.cpp:1712-1749) and
                     visited in block order, advected[p] += w * U_pre(corner) q,
                     then q' = sum_p After_p advected[p] in key-cell order.
training_column():   CUBATURE_GENERATOR_EIGEN::getStagedTrainingColumn
                     (src/cubature/CUBATURE_GENERATOR_EIGEN.cpp:300-350) with
                     SUBSPACE_FLUID_3D_EIGEN::advectCellStamPeeled
                     (SUBSPACE_FLUID_3D_EIGEN.cpp:910-1005): for every training
                     sample s, U_post(x)^T times the velocity advected to cell x,
                     scaled by the sample's magnitude factor.
magnitudes():        generateImportanceSampledCubature (:588-603): 1/sqrt(|b_s|^2)
                     evaluated in single precision, 1 when |b_s|^2 <= 1e-7.

U_pre(x) / U_after(x) are rows of the bases decoded in FP64 by the codec
oracle (DecodeVectorField, whose values GetSubmatrixNoSVD[Sparse] returns
cell by cell).  Pins (tests/test_cubature_oracle.py): with every interior cell
a key cell of weight 1 the cubature equals the exact reduced advection
U^T A(U q) (closure_oracle.reduced_advect, pinned by closed forms); the sum of
the training columns of every cell equals the stacked post-advection samples
when those are the exact reduced advection of the pre samples.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def dense_rows(basis, threads: int = 8) -> np.ndarray:
    """All columns of a basis decoded in FP64, as rows[cell, comp, col]."""
    X, Y, Z = basis.dims
    cols = basis.decode_vector_fields_mt(0, basis.r, threads)  # [r][3 * cells]
    return np.ascontiguousarray(cols.reshape(basis.r, X * Y * Z, 3).transpose(1, 2, 0))


def after_matrices(rows_after: np.ndarray, cells, weights) -> np.ndarray:
    """[K][r][3]: After_i = w_i * U_after(x_i)^T."""
    return np.stack([w * rows_after[c].T for c, w in zip(cells, weights)])


def _backtrace(dims, index, v, dt0):
    """accumAdvectRequests :1869-1934: peeled index + velocity -> the 8
    (peeled corner cell, weight) pairs of the clamped backtrace."""
    Xp, Yp, Zp = dims
    slab = Xp * Yp
    z = index // slab + 1
    y = (index % slab) // Xp + 1
    x = (index % slab) % Xp + 1
    xres, yres, zres = Xp + 2, Yp + 2, Zp + 2
    xt, yt, zt = x - dt0 * v[0], y - dt0 * v[1], z - dt0 * v[2]
    xt = 1.5 if xt < 1.5 else xt
    xt = xres - 2.5 if xt > xres - 2.5 else xt
    yt = 1.5 if yt < 1.5 else yt
    yt = yres - 2.5 if yt > yres - 2.5 else yt
    zt = 1.5 if zt < 1.5 else zt
    zt = zres - 2.5 if zt > zres - 2.5 else zt
    x0, y0, z0 = int(xt) - 1, int(yt) - 1, int(zt) - 1
    x1, y1, z1 = x0 + 1, y0 + 1, z0 + 1
    s1 = (xt - 1) - x0
    s0 = 1.0 - s1
    t1 = (yt - 1) - y0
    t0 = 1.0 - t1
    u1 = (zt - 1) - z0
    u0 = 1.0 - u1

    def cell(a, b, c):
        return a + b * Xp + c * slab

    return [(cell(x0, y0, z0), u0 * s0 * t0), (cell(x0, y1, z0), u0 * s0 * t1),
            (cell(x1, y0, z0), u0 * s1 * t0), (cell(x1, y1, z0), u0 * s1 * t1),
            (cell(x0, y0, z1), u1 * s0 * t0), (cell(x0, y1, z1), u1 * s0 * t1),
            (cell(x1, y0, z1), u1 * s1 * t0), (cell(x1, y1, z1), u1 * s1 * t1)]


def cubature_advect(rows_pre: np.ndarray, after: np.ndarray, cells, dims, q, dt0: float, return_advected=False):
    """reducedAdvectCompressionFriendly: q' of the cubature (FP64); with
    return_advected, also the interpolated velocity at every key cell."""
    q = np.asarray(q, np.float64)
    requests = []  # (block number, insertion order, key cell p, corner cell, weight)
    for p, index in enumerate(cells):
        v = rows_pre[index] @ q  # cellU * qDot (the 'before' rows of U_pre)
        for corner, w in _backtrace(dims, int(index), v, dt0):
            bn, _ = O.compute_block_number(3 * corner, dims)
            requests.append((bn, len(requests), p, corner, w))
    requests.sort(key=lambda t: (t[0], t[1]))  # multimap order: by block, then insertion
    advected = np.zeros((len(cells), 3))
    for _, _, p, corner, w in requests:
        advected[p] += w * (rows_pre[corner] @ q)
    final = np.zeros(after.shape[1])
    for p in range(len(cells)):
        final += after[p] @ advected[p]
    return (final, advected) if return_advected else final


def advect_cell(rows_pre: np.ndarray, dims, q, dt0: float, index: int) -> np.ndarray:
    """SUBSPACE_FLUID_3D_EIGEN::advectCellStamPeeled: U_pre q interpolated at
    the clamped backtrace of cell `index`, as the weighted sum of the eight
    corner values w000 v000 + ... + w111 v111 that it returns (:1001-1002)."""
    v = rows_pre[index] @ q
    c = _backtrace(dims, index, v, dt0)
    vals = [rows_pre[k] @ q for k, _ in c]
    (_, w000), (_, w010), (_, w100), (_, w110), (_, w001), (_, w011), (_, w101), (_, w111) = c
    return (w000 * vals[0] + w010 * vals[1] + w100 * vals[2] + w110 * vals[3] +
            w001 * vals[4] + w011 * vals[5] + w101 * vals[6] + w111 * vals[7])


def magnitudes(post: np.ndarray) -> np.ndarray:
    """_postadvectionMagnitudes: 1.0f / sqrtf(|b_s|^2) in single precision."""
    out = np.empty(len(post))
    for i, b in enumerate(np.asarray(post, np.float64)):
        m = float(b @ b)
        out[i] = float(np.float32(1.0) / np.sqrt(np.float32(m))) if abs(m) > 1e-7 else 1.0
    return out


def training_column(rows_pre, rows_post, dims, dt0, pre_samples, mags, index) -> np.ndarray:
    """getStagedTrainingColumn: stacked U_post(x)^T advectCell(q_s) * mag_s."""
    out = []
    for q, m in zip(pre_samples, mags):
        post = advect_cell(rows_pre, dims, np.asarray(q, np.float64), dt0, int(index))
        out.append((rows_post[index].T @ post) * m)
    return np.concatenate(out)
