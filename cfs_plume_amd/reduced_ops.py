"""The reduced per-step matrices of stepPlume, read from the reference's files.

initOutOfCore (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:84-139) requires, in the
reduced path, projected.ptof.matrix (P, pre-projection to final),
projected.vtod.matrix, projected.ptov.matrix, projected.A.matrix,
damping.peeled.matrix (D) and inverseProduct.matrix, and exits ("run
./buildProducts first") when any is missing.  When inverseProduct.matrix
exists but cannot be read, it is rebuilt as ptov * A^-1 * vtod and written
back.  Between the advection and the unprojection stepPlume (:453-501)
applies

    reducedPeeledDiffusion  (:848-854)   q <- D q
    reducedStagedProject    (:942-946)   q <- P q + inverseProduct q

All are r x r FP64 products on the host (at r = 200, 80 K multiply-adds);
q moves between host and device once per step.
"""
from __future__ import annotations

import os

import numpy as np

from .matrix_io import read_matrix, write_matrix

__all__ = ["ReducedOperators"]

_FILES = {
    "ptof": "projected.ptof.matrix",
    "vtod": "projected.vtod.matrix",
    "ptov": "projected.ptov.matrix",
    "A": "projected.A.matrix",
    "damping": "damping.peeled.matrix",
}
_INVERSE = "inverseProduct.matrix"


class ReducedOperators:
    def __init__(self, ptof: np.ndarray, damping: np.ndarray, inverse_product: np.ndarray):
        self.ptof = np.asarray(ptof, np.float64)
        self.damping = np.asarray(damping, np.float64)
        self.inverse_product = np.asarray(inverse_product, np.float64)
        r = self.ptof.shape[0]
        for name, M in (("ptof", self.ptof), ("damping", self.damping), ("inverseProduct", self.inverse_product)):
            if M.shape != (r, r):
                raise ValueError(f"{name} is {M.shape}, expected {r} x {r}")
        self.r = r

    @classmethod
    def load(cls, reduced_path: str) -> "ReducedOperators":
        """Reads the six matrices as initOutOfCore does (:84-139).  Any missing
        file raises FileNotFoundError (the reference prints "run ./buildProducts
        first" and exits); an inverseProduct.matrix that exists but does not
        read as a matrix is rebuilt from ptov, A and vtod and written back."""
        paths = {k: os.path.join(reduced_path, f) for k, f in _FILES.items()}
        inv_path = os.path.join(reduced_path, _INVERSE)
        missing = [p for p in [*paths.values(), inv_path] if not os.path.exists(p)]
        if missing:
            raise FileNotFoundError(f"pre-built reduced matrices missing (build the products first): {missing}")
        M = {k: read_matrix(p) for k, p in paths.items()}
        try:
            inverse_product = read_matrix(inv_path)
        except ValueError:
            inverse_product = M["ptov"] @ np.linalg.inv(M["A"]) @ M["vtod"]
            write_matrix(inv_path, inverse_product)
        return cls(M["ptof"], M["damping"], inverse_product)

    def diffuse(self, q: np.ndarray) -> np.ndarray:
        """reducedPeeledDiffusion: D q."""
        return self.damping @ np.asarray(q, np.float64)

    def pressure_project(self, q: np.ndarray) -> np.ndarray:
        """reducedStagedProject: P q + inverseProduct q."""
        q = np.asarray(q, np.float64)
        return self.ptof @ q + self.inverse_product @ q

    def after_advection(self, q: np.ndarray) -> np.ndarray:
        """The two stages in stepPlume's order: diffusion, then pressure projection."""
        return self.pressure_project(self.diffuse(q))
