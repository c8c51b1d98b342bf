"""Dense `.matrix` files in the reference's layout.

EIGEN::read / EIGEN::write (src/linearalgebra/EIGEN.cpp:24-121, 245-290):
int32 rows, int32 cols, then rows x cols FP64 in row-major order, native
(little-endian) byte order.  A name ending in "gz" is refused (the reference
exits there), and so are a short header, non-positive dimensions and a short
payload (where the reference returns false).

The per-step subspace trajectories of the accuracy harness use the same
layout, written as the dimensions first and one FP64 row appended per step
(SUBSPACE_FLUID_3D_EIGEN::writeSubspaceErrorMatrixDims / appendSubspaceVectors,
src/integrators/SUBSPACE_FLUID_3D_EIGEN.cpp:1796-1840).
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["read_matrix", "write_matrix", "TrajectoryWriter"]


def _check_name(path) -> str:
    p = os.fspath(path)
    if p.endswith("gz"):
        raise ValueError(f"{p}: .matrix files are not gzipped (EIGEN::read/write refuse such names)")
    return p


def write_matrix(path, M: np.ndarray) -> None:
    """EIGEN::write: a 2-D array as int32 rows, int32 cols, FP64 row-major."""
    p = _check_name(path)
    A = np.ascontiguousarray(M, dtype="<f8")
    if A.ndim != 2 or A.shape[0] <= 0 or A.shape[1] <= 0:
        raise ValueError(f"write_matrix needs a non-empty 2-D array, got shape {A.shape}")
    with open(p, "wb") as f:
        f.write(np.array(A.shape, dtype="<i4").tobytes())
        f.write(A.tobytes())


def read_matrix(path) -> np.ndarray:
    """EIGEN::read: the FP64 [rows][cols] array of a `.matrix` file."""
    p = _check_name(path)
    with open(p, "rb") as f:
        hdr = f.read(8)
        if len(hdr) != 8:
            raise ValueError(f"{p}: short header")
        rows, cols = (int(x) for x in np.frombuffer(hdr, dtype="<i4"))
        if rows <= 0 or cols <= 0:
            raise ValueError(f"{p}: bad dimensions {rows} x {cols}")
        data = np.fromfile(f, dtype="<f8", count=rows * cols)
    if data.size != rows * cols:
        raise ValueError(f"{p}: {data.size} of {rows * cols} values")
    return data.reshape(rows, cols)


class TrajectoryWriter:
    """A [steps][r] trajectory written one row at a time: the dimensions at
    construction, then append(q) once per step."""

    def __init__(self, path, steps: int, r: int):
        if steps <= 0 or r <= 0:
            raise ValueError("steps and r must be positive")
        self.path = _check_name(path)
        self.r = int(r)
        with open(self.path, "wb") as f:
            f.write(np.array([steps, r], dtype="<i4").tobytes())

    def append(self, q: np.ndarray) -> None:
        v = np.ascontiguousarray(q, dtype="<f8").reshape(-1)
        if v.size != self.r:
            raise ValueError(f"row of {v.size} values for r = {self.r}")
        with open(self.path, "ab") as f:
            f.write(v.tobytes())
