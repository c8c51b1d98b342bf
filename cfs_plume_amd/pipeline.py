"""The plume data pipeline: full-rank snapshots -> SVD bases -> reduced
operators -> cubature -> compressed bases, on the device.

Reference (aadjones/CFS_PLUME) -> here:

  projects/fluid3D (FLUID_3D_MIC step + appendStreams)   fluid.FluidSolver.run
  projects/svdOutOfCoreMultiple (BIG_MATRIX::outOfCoreSVD,
      BIG_MATRIX.cpp:276-420; keep rule :326-332)          svd_basis
  projects/buildProducts (buildOutOfCoreMatrices,
      SUBSPACE_FLUID_3D_EIGEN.cpp:1355-1489)               build_products
  projects/cubatureGeneratorStamStaged (training pairs
      cubatureGeneratorStamStaged.cpp:150-200)             training_pairs + cubature.generate_cubature
  projects/compressMatrices                              basis.encode_matrix (compress_bases)

The snapshot matrices, bases and operator products stay on the device in
FP64 (column-major: one snapshot or mode per column, i.e. a DeviceArray of
shape (columns, rows)); the small eigenproblems (snapshot count squared) and
the r x r inverse run on the host.  The SVD is the method of snapshots:
G = A^T A on the device, G = V L V^T on the host, U = A V L^-1/2 on the
device, then one re-orthonormalisation (U <- U R^-1 with R the Cholesky
factor of U^T U); its left singular vectors and PCA values L are those of
the reference's QR-then-SVD (BIG_MATRIX.cpp:276-420), up to each column's
sign.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _lib
from .basis import CompressedBasis, _sptr, encode_matrix
from .cubature import CubatureResult, generate_cubature
from .fluid import FluidSolver, Snapshots
from .hip import DeviceArray, Stream
from .matrix_io import write_matrix

OP_POISSON, OP_VEL_TO_DIV, OP_PRESSURE_TO_VEL, OP_DAMPING = 0, 1, 2, 3
STREAM_BASES = {"final_velocity": "final", "preproject": "preproject", "prediffuse": "prediffuse",
                "preadvect": "preadvect"}


def gemm_tn(A: DeviceArray, B: DeviceArray, stream: Optional[Stream] = None) -> np.ndarray:
    """A^T B for device matrices of shapes (m, rows) and (n, rows): (m, n) on the host."""
    m, rows = A.shape
    n, rows_b = B.shape
    if rows != rows_b:
        raise ValueError(f"row counts differ: {rows} and {rows_b}")
    C = DeviceArray((m, n), np.float64)
    _lib.check(_lib.lib().cfs_gemm_tn(A.ptr, rows, m, rows, B.ptr, n, rows, C.ptr, _sptr(stream)), "gemm_tn")
    return C.numpy()


def gemm_nn(A: DeviceArray, M: np.ndarray, stream: Optional[Stream] = None) -> DeviceArray:
    """A M for a device matrix A of shape (k, rows) and a host k x n M: shape (n, rows)."""
    k, rows = A.shape
    M = np.ascontiguousarray(M, np.float64)
    if M.shape[0] != k:
        raise ValueError(f"inner dimensions differ: {k} and {M.shape[0]}")
    n = M.shape[1]
    dM = DeviceArray.from_numpy(M)
    C = DeviceArray((n, rows), np.float64)
    _lib.check(_lib.lib().cfs_gemm_nn(A.ptr, rows, k, rows, dM.ptr, n, C.ptr, rows, _sptr(stream)), "gemm_nn")
    return C


def apply_operator(op: int, full, boundaries, dx: float, w: float, V: DeviceArray,
                   stream: Optional[Stream] = None) -> DeviceArray:
    """The sparse operator `op` on every column of V (shape (k, rows_in))."""
    X, Y, Z = full
    np_ = (X - 2) * (Y - 2) * (Z - 2)
    rows_in = np_ if op in (OP_POISSON, OP_PRESSURE_TO_VEL) else 3 * np_
    rows_out = np_ if op in (OP_POISSON, OP_VEL_TO_DIV) else 3 * np_
    k, rows = V.shape
    if rows != rows_in:
        raise ValueError(f"operator {op} takes columns of {rows_in} values, got {rows}")
    out = DeviceArray((k, rows_out), np.float64)
    bc = (ctypes.c_int * 6)(*boundaries)
    _lib.check(_lib.lib().cfs_apply_operator(op, X, Y, Z, bc, float(dx), float(w), V.ptr, rows_in, k, out.ptr,
                                             rows_out, _sptr(stream)), "apply operator")
    return out


@dataclass
class Basis:
    U: DeviceArray       # (r, rows): mode j is row j
    pca: np.ndarray      # the PCA values (squared singular values) of all kept snapshots, decreasing
    snapshots: int       # snapshots used (norm above the floor)

    @property
    def r(self) -> int:
        return self.U.shape[0]

    def numpy(self) -> np.ndarray:
        return self.U.numpy()


def svd_basis(snaps: DeviceArray, threshold: float = 1e-9, norm_floor: float = 1e-7,
              stream: Optional[Stream] = None) -> Basis:
    """The left singular vectors of the snapshot matrix (columns = the rows of
    `snaps`, shape (steps, rows)) whose PCA values s_i^2 satisfy
    s_i^2 / s_0^2 >= threshold (outOfCoreSVD's keep rule); snapshots with
    norm <= norm_floor are left out, as appendStreams never writes them.
    Column signs: the largest-magnitude entry of each mode is positive."""
    G = gemm_tn(snaps, snaps, stream)
    keep = np.sqrt(np.maximum(np.diag(G), 0.0)) > norm_floor
    idx = np.flatnonzero(keep)
    if idx.size == 0:
        raise ValueError("every snapshot is below the norm floor")
    lam, V = np.linalg.eigh(G[np.ix_(idx, idx)])
    order = np.argsort(lam)[::-1]
    lam, V = lam[order], V[:, order]
    k = 0
    while k < lam.size and lam[k] / lam[0] >= threshold:
        k += 1
    M = np.zeros((snaps.shape[0], k))
    M[idx] = V[:, :k] / np.sqrt(lam[:k])
    U = gemm_nn(snaps, M, stream)
    # re-orthonormalise: U <- U R^-1, R^T R = U^T U (close to the identity)
    R = np.linalg.cholesky(gemm_tn(U, U, stream)).T
    U2 = gemm_nn(U, np.linalg.inv(R), stream)
    host = U2.numpy()
    signs = np.sign(host[np.arange(k), np.argmax(np.abs(host), axis=1)])
    if np.any(signs < 0):
        U2.upload(host * signs[:, None])
    return Basis(U2, lam, int(idx.size))


def build_products(full, boundaries, dx: float, bases: dict, w: float = 0.9,
                   stream: Optional[Stream] = None, iop: bool = False) -> dict:
    """buildOutOfCoreMatrices: the reduced operators from the bases
    (pressure, preproject, prediffuse, final) and the sparse full-grid
    operators, FP64 on the host (names as the reference's files).

    iop: buildOutOfCoreMatricesIOP (SUBSPACE_FLUID_3D_EIGEN.cpp:1494-1659):
    the pressure projection acts on the post-stomp velocity, so vtod and ptof
    take U_iop (bases["iop"]) in place of U_preproject, and U_iop^T
    U_preproject ("U'U.iop.subspace", :1635-1637) is added."""
    Up, Uproj, Udiff, Ufin = (bases[k].U for k in ("pressure", "preproject", "prediffuse", "final"))
    Uvel = bases["iop"].U if iop else Uproj
    A = gemm_tn(Up, apply_operator(OP_POISSON, full, boundaries, dx, w, Up, stream), stream)
    vtod = gemm_tn(Up, apply_operator(OP_VEL_TO_DIV, full, boundaries, dx, w, Uvel, stream), stream)
    damping = gemm_tn(Uproj, apply_operator(OP_DAMPING, full, boundaries, dx, w, Udiff, stream), stream)
    ptof = gemm_tn(Ufin, Uvel, stream)
    ptov = gemm_tn(Ufin, apply_operator(OP_PRESSURE_TO_VEL, full, boundaries, dx, w, Up, stream), stream)
    inverse = ptov @ np.linalg.inv(A) @ vtod
    out = {"projected.A": A, "projected.vtod": vtod, "damping.peeled": damping, "projected.ptof": ptof,
           "projected.ptov": ptov, "inverseProduct": inverse}
    if iop:
        out["U'U.iop.subspace"] = gemm_tn(bases["iop"].U, Uproj, stream)
    return out


def training_pairs(snaps: Snapshots, bases: dict, count: int) -> tuple:
    """The cubature generator's training samples (cubatureGeneratorStamStaged.cpp:
    150-200): for the first `count` snapshots, q_pre = U_preadvect^T v_preadvect
    and q_post = U_prediffuse^T v_prediffuse of the full-rank run."""
    pre = bases["preadvect"].U
    post = bases["prediffuse"].U
    a = snaps.arrays["preadvect"]
    b = snaps.arrays["prediffuse"]
    n = min(count, snaps.steps)
    sub_a = DeviceArray.view(a.ptr, (n, a.shape[1]), np.float64, a)
    sub_b = DeviceArray.view(b.ptr, (n, b.shape[1]), np.float64, b)
    return gemm_tn(sub_a, pre).copy(), gemm_tn(sub_b, post).copy()


def dense_rows_of(U: np.ndarray, dims) -> np.ndarray:
    """rows[cell, comp, mode] of a dense basis (r, 3 Npeel)."""
    X, Y, Z = dims
    return np.ascontiguousarray(U.reshape(U.shape[0], X * Y * Z, 3).transpose(1, 2, 0))


@dataclass
class PlumeData:
    full: tuple
    snapshots: Snapshots
    bases: dict
    operators: dict
    cubature: Optional[CubatureResult] = None
    compressed: dict = field(default_factory=dict)
    timings: dict = field(default_factory=dict)
    box: Optional[object] = None  # the obstacle's box at the run's start (moving-obstacle data)

    @property
    def peeled(self):
        return tuple(d - 2 for d in self.full)


def build_plume_data(full=(48, 64, 48), steps: int = 100, threshold: float = 1e-9, cubature_snapshots: int = 20,
                     candidates: int = 1000, tolerance: float = 0.01, passes: int = 6, percent: float = 0.9995,
                     n_bits: int = 32, max_iterations: int = 50, threads: int = 16, vorticity: float = 0.0,
                     obstacle=None, log: Optional[Callable[[str], None]] = None) -> PlumeData:
    """The reference's offline chain for one plume (cfg stam.64: 48 x 64 x 48,
    100 snapshots, percent 0.9995, nBits 32, maxIterations 50; cubature from
    the first 20 snapshot pairs, 1000 candidates per pass, error 0.01).

    obstacle: an obstacle.Box -- the moving-obstacle chain instead: the
    full-rank run steps with the box (stepWithMovingObstacle, no buoyancy,
    passive smoke), the IOP stream gets its basis U.iop, and the products
    are the IOP ones (build_products(iop=True)); the box is kept at its
    starting position in PlumeData.box."""
    import time
    say = log or (lambda m: None)
    t = {}
    t0 = time.perf_counter()
    fluid = FluidSolver(full, vorticity=vorticity)
    start_box = obstacle.copy() if obstacle is not None else None
    snaps = fluid.run_obstacle(steps, obstacle.copy(), log=say) if obstacle is not None else fluid.run(steps, log=say)
    prm = fluid.params
    boundaries, dx = list(prm.boundaries), prm.dx
    dt0 = prm.dt / prm.dx
    t["full_rank_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    bases = {}
    stream_bases = dict(STREAM_BASES, **({"iop": "iop"} if obstacle is not None else {}))
    for stream_name, name in stream_bases.items():
        bases[name] = svd_basis(snaps.arrays[stream_name], threshold)
        say(f"U.{name}: {bases[name].r} modes of {bases[name].snapshots} snapshots")
    # the pressure stream holds the pressure and the divergence of every step
    p, d = snaps.arrays["pressure"], snaps.arrays["divergence"]
    both = DeviceArray((2 * steps, p.shape[1]), np.float64)
    both.upload(np.stack([p.numpy(), d.numpy()], axis=1).reshape(2 * steps, -1))
    bases["pressure"] = svd_basis(both, threshold)
    say(f"U.pressure: {bases['pressure'].r} modes of {bases['pressure'].snapshots} snapshots")
    t["svd_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ops = build_products(full, boundaries, dx, bases, iop=obstacle is not None)
    t["products_s"] = time.perf_counter() - t0
    data = PlumeData(tuple(full), snaps, bases, ops, timings=t, box=start_box)
    # cubature
    t0 = time.perf_counter()
    q_pre, q_post = training_pairs(snaps, bases, cubature_snapshots)
    dims = data.peeled
    rows_pre = dense_rows_of(bases["preadvect"].numpy(), dims)
    rows_post = dense_rows_of(bases["prediffuse"].numpy(), dims)
    data.cubature = generate_cubature(rows_pre, rows_post, dims, q_pre, q_post, dt0,
                                      candidates_per_pass=candidates, passes=passes, tolerance=tolerance,
                                      log=lambda m: say(m))
    t["cubature_s"] = time.perf_counter() - t0
    # compressed U_preadvect and U_final
    t0 = time.perf_counter()
    for name in ("preadvect", "final"):
        data.compressed[name] = encode_matrix(bases[name].numpy(), dims, n_bits, percent, max_iterations,
                                              threads=threads).buffers()
    t["compress_s"] = time.perf_counter() - t0
    fluid.close()
    return data


def write_reduced(data: PlumeData, path: str) -> None:
    """The reduced path's files: the six operator matrices, the cubature,
    the PCA values and the compressed U.preadvect / U.final components
    (the names the reference's solvers read)."""
    from .cubature import write_cubature
    os.makedirs(path, exist_ok=True)
    for name, M in data.operators.items():
        write_matrix(os.path.join(path, name + ".matrix"), M)
    if data.cubature is not None:
        write_cubature(os.path.join(path, "cubature"), data.cubature.cells, data.cubature.weights)
    for name, bufs in data.compressed.items():
        for k, b in zip("XYZ", bufs):
            with open(os.path.join(path, f"U.{name}.component{k}"), "wb") as f:
                f.write(b)


def compressed_bases(data: PlumeData) -> dict:
    return {name: CompressedBasis.from_buffers(bufs) for name, bufs in data.compressed.items()}
