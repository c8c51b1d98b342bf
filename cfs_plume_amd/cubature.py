"""Cubature reduced advection and its generator (host side over the C ABI).

Reference (aadjones/CFS_PLUME) -> here:

  readAdvectionCubature (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:     Cubature.from_file / from_cells
    1188-1327), the `cubature` file (CUBATURE_GENERATOR_EIGEN.cpp:50-91)  (cfs_cubature_open_file / create),
                                                                      write_cubature
  reducedAdvectCompressionFriendly (:1781-1854)                      Cubature.advect (cfs_cubature_advect)
  CUBATURE_GENERATOR_EIGEN::generateImportanceSampledCubature         generate_cubature
    (:567-688), pickImportanceSampledCandidatesOMP (:96-200),
    getStagedTrainingColumn (:300-350), cullCubatureZeros (:543-563)
  NNLS_SOLVER::solve (NNLS.h:156-208)                                nnls (cfs_nnls)

The generator is offline host code, as in the reference: training columns
are evaluated with numpy on the dense decoded bases (rows[cell, comp, col]),
the non-negative fits by the library's Lawson-Hanson solver.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from .basis import CompressedBasis, _sptr
from .hip import DeviceArray, Stream


def _p(a):
    """Device pointer of a DeviceArray or (DeviceArray, row) into an [n][r] array."""
    if a is None:
        return None
    if isinstance(a, tuple):
        arr, row = a
        return arr.ptr + row * arr.shape[-1] * arr.dtype.itemsize
    return a.ptr


class Cubature:
    """Key cells, weights and after-matrices on the device."""

    def __init__(self, handle):
        self._h = handle
        K, r = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().cfs_cubature_get_info(self._h, ctypes.byref(K), ctypes.byref(r)), "cubature info")
        self.K, self.r = K.value, r.value

    @classmethod
    def from_cells(cls, cells, weights, after: CompressedBasis) -> "Cubature":
        c = np.ascontiguousarray(cells, np.int32)
        w = np.ascontiguousarray(weights, np.float64)
        if c.shape != w.shape or c.ndim != 1:
            raise ValueError("cells and weights must be 1-D arrays of one length")
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_cubature_create(len(c), c.ctypes.data, w.ctypes.data, after._h, ctypes.byref(h)),
                   "readAdvectionCubature")
        return cls(h.value)

    @classmethod
    def from_rows(cls, cells, weights, dims, U_after: np.ndarray) -> "Cubature":
        """With a dense after-basis (U.prediffuse): U_after is r x 3 Npeel (one
        mode per row, components interleaved) on the host; dims the peeled grid."""
        c = np.ascontiguousarray(cells, np.int32)
        w = np.ascontiguousarray(weights, np.float64)
        if c.shape != w.shape or c.ndim != 1:
            raise ValueError("cells and weights must be 1-D arrays of one length")
        U = np.asarray(U_after, np.float64)
        r = U.shape[0]
        # entry (comp, j) of cell i at i 3r + comp + 3j
        rows = np.ascontiguousarray(U.reshape(r, -1, 3)[:, c, :].transpose(1, 0, 2))
        h = ctypes.c_void_p()
        X, Y, Z = dims
        _lib.check(_lib.lib().cfs_cubature_create_rows(len(c), c.ctypes.data, w.ctypes.data, X, Y, Z, r,
                                                       rows.ctypes.data, ctypes.byref(h)), "readAdvectionCubature")
        return cls(h.value)

    @classmethod
    def from_file(cls, path: str, after: CompressedBasis) -> "Cubature":
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_cubature_open_file(path.encode(), after._h, ctypes.byref(h)), "readAdvectionCubature")
        return cls(h.value)

    def cells_and_weights(self):
        c = np.empty(self.K, np.int32)
        w = np.empty(self.K, np.float64)
        _lib.check(_lib.lib().cfs_cubature_get_cells(self._h, c.ctypes.data, w.ctypes.data), "cubature cells")
        return c, w

    def advect(self, pre: CompressedBasis, q_in, q_out, scratch: DeviceArray, dt0: float, alpha: float = 1.0,
               forcing=None, stream: Optional[Stream] = None):
        """q_out = alpha * reducedAdvectCompressionFriendly(q_in) + forcing;
        scratch is a full-grid FP32 field (interior overwritten)."""
        _lib.check(_lib.lib().cfs_cubature_advect(self._h, pre._h, _p(q_in), _p(q_out), scratch.ptr, float(dt0),
                                                  float(alpha), _p(forcing), _sptr(stream)),
                   "reducedAdvectCompressionFriendly")

    def close(self):
        if self._h:
            _lib.lib().cfs_cubature_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_cubature(path: str, cells, weights) -> None:
    """writeCubature: int32 K, int32 cells[K], FP64 weights[K]."""
    c = np.ascontiguousarray(cells, np.int32)
    w = np.ascontiguousarray(weights, np.float64)
    _lib.check(_lib.lib().cfs_cubature_write_file(path.encode(), len(c), c.ctypes.data, w.ctypes.data),
               "writeCubature")


def nnls(A: np.ndarray, b: np.ndarray, max_iter_factor: int = 5):
    """min |A x - b| subject to x >= 0 (Lawson-Hanson): (x, residual norm,
    inner iterations, capped)."""
    Af = np.asfortranarray(A, dtype=np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    m, n = Af.shape
    if bb.shape != (m,):
        raise ValueError("b must have one entry per row of A")
    x = np.zeros(n)
    rn, it = ctypes.c_double(), ctypes.c_int()
    _lib.check(_lib.lib().cfs_nnls(Af.ctypes.data, m, n, bb.ctypes.data, max_iter_factor, x.ctypes.data,
                                   ctypes.byref(rn), ctypes.byref(it)), "NNLS")
    return x, rn.value, it.value, it.value > max_iter_factor * n


class MersenneTwister:
    """MT19937 with init_genrand seeding and the draws the generator uses:
    randInt(n) uniform on [0, n] by masked rejection, rand() = randInt() /
    (2^32 - 1) (src/util/MERSENNETWISTER.h:146-147, 179-205)."""

    def __init__(self, seed: int):
        self._bg = np.random.MT19937()
        self._bg._legacy_seeding(int(seed) & 0xFFFFFFFF)
        self._buf = np.empty(0, np.uint64)
        self._i = 0

    def rand_int32(self) -> int:
        if self._i >= len(self._buf):
            self._buf = self._bg.random_raw(4096)
            self._i = 0
        v = int(self._buf[self._i])
        self._i += 1
        return v

    def rand_int(self, n: int) -> int:
        used = n
        for s in (1, 2, 4, 8, 16):
            used |= used >> s
        while True:
            i = self.rand_int32() & used
            if i <= n:
                return i

    def rand(self) -> float:
        return self.rand_int32() * (1.0 / 4294967295.0)


def advected_fields(rows_pre: np.ndarray, dims, q_samples: np.ndarray, dt0: float) -> np.ndarray:
    """[S][cells][3]: for every sample, U_pre q advected to every peeled cell
    (advectCellStamPeeled for all cells: backtrace by U_pre(x) q, clamp to
    [1.5, res - 2.5], weighted sum of the 8 corner values)."""
    X, Y, Z = dims
    n = X * Y * Z
    idx = np.arange(n)
    z = idx // (X * Y) + 1
    y = (idx % (X * Y)) // X + 1
    x = (idx % (X * Y)) % X + 1
    out = np.empty((len(q_samples), n, 3))
    for s, q in enumerate(np.asarray(q_samples, np.float64)):
        V = rows_pre @ q  # [cells][3]
        xt = np.clip(x - dt0 * V[:, 0], 1.5, X + 2 - 2.5)
        yt = np.clip(y - dt0 * V[:, 1], 1.5, Y + 2 - 2.5)
        zt = np.clip(z - dt0 * V[:, 2], 1.5, Z + 2 - 2.5)
        x0, y0, z0 = xt.astype(np.int64) - 1, yt.astype(np.int64) - 1, zt.astype(np.int64) - 1
        s1, t1, u1 = (xt - 1) - x0, (yt - 1) - y0, (zt - 1) - z0
        s0, t0, u0 = 1.0 - s1, 1.0 - t1, 1.0 - u1

        def at(dx, dy, dz):
            return V[(x0 + dx) + (y0 + dy) * X + (z0 + dz) * X * Y]

        out[s] = ((u0 * s0 * t0)[:, None] * at(0, 0, 0) + (u0 * s0 * t1)[:, None] * at(0, 1, 0) +
                  (u0 * s1 * t0)[:, None] * at(1, 0, 0) + (u0 * s1 * t1)[:, None] * at(1, 1, 0) +
                  (u1 * s0 * t0)[:, None] * at(0, 0, 1) + (u1 * s0 * t1)[:, None] * at(0, 1, 1) +
                  (u1 * s1 * t0)[:, None] * at(1, 0, 1) + (u1 * s1 * t1)[:, None] * at(1, 1, 1))
    return out


def magnitudes(post: np.ndarray) -> np.ndarray:
    """1.0f / sqrtf(|b_s|^2) in single precision, 1 when |b_s|^2 <= 1e-7
    (generateImportanceSampledCubature :588-603)."""
    m = np.einsum("sr,sr->s", post, post)
    with np.errstate(divide="ignore"):  # zero samples (a fluid at rest) take 1 below
        inv = (np.float32(1.0) / np.sqrt(m.astype(np.float32))).astype(np.float64)
    return np.where(np.abs(m) > 1e-7, inv, 1.0)


@dataclass
class CubatureResult:
    cells: np.ndarray
    weights: np.ndarray
    relative_error: float
    passes: int
    history: list = field(default_factory=list)


def generate_cubature(rows_pre: np.ndarray, rows_post: np.ndarray, dims, pre: np.ndarray, post: np.ndarray,
                      dt0: float, candidates_per_pass: int = 5000, passes: int = 6, tolerance: float = 0.01,
                      nnls_max_iter: int = 5, seed: int = 123456,
                      log: Optional[Callable[[str], None]] = None) -> CubatureResult:
    """generateImportanceSampledCubature: key cells and non-negative weights
    such that sum_i w_i g_i fits the stacked, magnitude-normalised
    post-advection samples b, where g_i is cell i's training column.

    Each pass draws candidates_per_pass new candidates by importance
    sampling (a uniformly drawn cell is accepted when a uniform draw falls
    below |g_i . residual| / (|residual|^2 / cells)), refits the weights of all
    candidates kept so far by NNLS, drops zero weights and stops once the
    relative residual is below `tolerance`.  As in the reference, the set of
    cells already used is local to each pass (the map is passed by value),
    so a cell may be drawn again in a later pass; the draw of cell indices
    uses a twister seeded pass * 314159, the accept draws one seeded `seed`
    that lives through all passes.  The sampling probabilities come from
    every cell's column dotted with the residual at once, which is the same
    number the reference computes per draw."""
    X, Y, Z = dims
    n = X * Y * Z
    pre = np.asarray(pre, np.float64)
    post = np.asarray(post, np.float64)
    S, r_post = post.shape
    if candidates_per_pass > n:
        raise ValueError(f"{candidates_per_pass} candidates per pass but the grid has {n} cells")
    mags = magnitudes(post)
    b = (post * mags[:, None]).reshape(-1)
    adv = advected_fields(rows_pre, dims, pre, dt0)  # [S][cells][3]

    def columns(cells):
        # g_i[s * r + j] = mags[s] * sum_c U_post[i, c, j] * adv[s, i, c]
        g = np.einsum("sic,icj->isj", adv[:, cells, :], rows_post[cells]) * mags[None, :, None]
        return g.reshape(len(cells), S * r_post).T

    def dots(residual):
        # every cell's column dotted with the residual: sum_s mags[s] adv[s, i] . (U_post[i] res_s)
        res = residual.reshape(S, r_post) * mags[:, None]
        W = np.einsum("icj,sj->sic", rows_post, res)
        return np.einsum("sic,sic->i", adv, W)

    accept_twister = MersenneTwister(seed)
    residual = b.copy()
    candidates: list = []
    weights = np.zeros(0)
    rel = 1.0
    history = []
    done = 0
    for p in range(passes):
        # pickImportanceSampledCandidatesOMP (:96-200), one thread
        scale = (residual @ residual) / n
        prob = np.abs(dots(residual)) / scale
        draw = MersenneTwister(0 + p * 314159)
        used = np.zeros(n, bool)
        found = 0
        while found < candidates_per_pass:
            i = draw.rand_int(n - 1)
            if used[i]:
                continue
            if accept_twister.rand() >= prob[i]:
                continue
            candidates.append(i)
            used[i] = True
            found += 1
        A = columns(np.asarray(candidates))
        weights, _, iters, capped = nnls(A, b, nnls_max_iter)
        residual = b - A @ weights
        rel = float(np.linalg.norm(residual) / np.linalg.norm(b))
        # cullCubatureZeros (:543-563)
        keep = np.abs(weights) >= 1e-8
        candidates = [c for c, k in zip(candidates, keep) if k]
        weights = weights[keep]
        history.append({"pass": p, "candidates": int(keep.size), "kept": int(keep.sum()), "rel_error": rel,
                        "nnls_iterations": iters, "nnls_capped": bool(capped)})
        if log:
            log(f"cubature pass {p}: {keep.size} candidates, {keep.sum()} kept, relative error {rel:.4e}")
        done = p + 1
        if rel < tolerance:
            break
    return CubatureResult(np.asarray(candidates, np.int32), np.asarray(weights, np.float64), rel, done, history)


def dense_rows(U: CompressedBasis, stream: Optional[Stream] = None) -> np.ndarray:
    """The basis decoded on the device (DecodeVectorField for every column),
    as FP64 rows[cell, comp, col] for the generator."""
    cols = U.decode_columns(0, U.r, None, stream).numpy()  # [r][3 * cells]
    X, Y, Z = U.dims
    return np.ascontiguousarray(cols.reshape(U.r, X * Y * Z, 3).transpose(1, 2, 0).astype(np.float64))
