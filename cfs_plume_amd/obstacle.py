"""The moving box obstacle (host side over the C ABI).

Reference (aadjones/CFS_PLUME) -> here:

  BOX (src/geometry/BOX.h, BOX.cpp:21-131)           Box (cfs_box_init / cfs_box_advance)
  projects/buildProducts/buildProducts.cpp:11-18     DEFAULT_* (the box of the obstacle runs)
  reducedSetMovingBox with buildObstacleMatrices'     reduced_iop (cfs_reduced_iop: the per-step
    pair (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:      R q + n evaluated on the device from the
    779-790, SUBSPACE_FLUID_3D_EIGEN.cpp:1699-1711)    box's cells, no per-step matrix files)

The full-rank obstacle step is FluidSolver.step_obstacle, the reduced one
PlumeSolver.step_obstacle.
"""
from __future__ import annotations

import copy
import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .basis import _sptr
from .hip import DeviceArray, Stream

DEFAULT_CENTER = (0.5, 0.5, 0.5)
DEFAULT_LENGTHS = (0.4, 0.05, 0.15)
DEFAULT_PERIOD = 4.0
DEFAULT_TRANSLATION_PERIOD = 10.0


class Box:
    """A box spinning about z (period) and sliding along x (translation
    period) in the unit cube; `state` is the C struct the kernels read."""

    def __init__(self, center=DEFAULT_CENTER, lengths=DEFAULT_LENGTHS, period: float = DEFAULT_PERIOD,
                 translation_period: float = DEFAULT_TRANSLATION_PERIOD, dt: float = 0.1):
        self.state = _lib.Box()
        c = (ctypes.c_double * 3)(*center)
        ln = (ctypes.c_double * 3)(*lengths)
        _lib.check(_lib.lib().cfs_box_init(c, ln, float(period), float(translation_period), float(dt),
                                           ctypes.byref(self.state)), "BOX")

    def advance(self):
        """translate_center, spin, update_time."""
        _lib.check(_lib.lib().cfs_box_advance(ctypes.byref(self.state)), "BOX advance")

    def copy(self) -> "Box":
        b = copy.copy(self)
        b.state = _lib.Box()
        ctypes.pointer(b.state)[0] = self.state
        return b

    @property
    def center(self):
        return np.array(self.state.center[:])

    @property
    def theta(self):
        return self.state.theta

    @property
    def phi(self):
        return self.state.phi

    @property
    def ref(self):
        return ctypes.byref(self.state)


def scratch_for(peeled_dims) -> DeviceArray:
    X, Y, Z = peeled_dims
    return DeviceArray((int(_lib.lib().cfs_reduced_iop_scratch_bytes(X, Y, Z)),), np.uint8)


def reduced_iop(U_iop: DeviceArray, U_proj: DeviceArray, M: DeviceArray, peeled_dims, box: Box, q: DeviceArray,
                out: Optional[DeviceArray] = None, scratch: Optional[DeviceArray] = None,
                stream: Optional[Stream] = None) -> DeviceArray:
    """q_out = R q + n for the box's current position.  U_iop, U_proj: device
    (r, 3 Npeel) FP64 (each row a column of the reference's dense basis);
    M = U_iop^T U_proj on the device (r_iop, r_proj)."""
    r_iop, rows = U_iop.shape
    r_proj = U_proj.shape[0]
    X, Y, Z = peeled_dims
    assert rows == 3 * X * Y * Z and U_proj.shape[1] == rows and M.shape == (r_iop, r_proj)
    assert q.dtype == np.float64 and q.nbytes == 8 * r_proj
    if out is None:
        out = DeviceArray.zeros((r_iop,), np.float64)
    _lib.check(_lib.lib().cfs_reduced_iop(U_iop.ptr, r_iop, U_proj.ptr, r_proj, M.ptr, X, Y, Z, box.ref, q.ptr,
                                          out.ptr, scratch.ptr if scratch is not None else None, _sptr(stream)),
               "reducedSetMovingBox")
    return out
