"""The whole compressed plume timestep on the device (host side over the C ABI).

Reference (aadjones/CFS_PLUME) -> here:

  FLUID_3D_MIC constants (FLUID_3D_MIC.cpp:30-56)            default_params
  addSmokeColumn (FLUID_3D_MIC.cpp:655-703)                 PlumeSolver.add_smoke (cfs_plume_add_smoke)
  stepPlume (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:453-501) PlumeSolver.step (cfs_plume_step)
  initOutOfCoreIOP (:145-200)                               PlumeSolver.set_obstacle (cfs_plume_set_obstacle)
  stepMovingObstacle (:508-566)                             PlumeSolver.step_obstacle (cfs_plume_step_obstacle)

A step is kernel launches on one stream (forces, project, density and heat
advection, subspace advection, reduced diffusion and pressure projection,
unproject), so it can be captured in a HIP graph (PlumeSolver.capture).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .basis import CompressedBasis, _sptr
from .hip import DeviceArray, Graph, Stream

ADVECT_EXACT, ADVECT_CUBATURE, ADVECT_CLOSURE = 0, 1, 2


def default_params(full_dims) -> _lib.PlumeParams:
    """dt 0.1, dx 1 / max resolution, buoyancy 0.1, vorticity 2 / max(1, 64 /
    max resolution); walls on x and z, open y (the plume driver's sides)."""
    p = _lib.PlumeParams()
    X, Y, Z = full_dims
    _lib.lib().cfs_plume_default_params(X, Y, Z, ctypes.byref(p))
    return p


class PlumeSolver:
    """stepPlume with U_preadvect, U_final and the reduced matrices on the device."""

    def __init__(self, pre: CompressedBasis, fin: CompressedBasis, damping: np.ndarray, ptof: np.ndarray,
                 inverse_product: np.ndarray, params: Optional[_lib.PlumeParams] = None):
        """damping: r_proj x r_adv; ptof, inverse_product: r_fin x r_proj (r_fin = fin.r)."""
        X, Y, Z = pre.dims
        self.full = (X + 2, Y + 2, Z + 2)
        self.r = pre.r
        self.params = params if params is not None else default_params(self.full)
        mats = [np.ascontiguousarray(m, np.float64) for m in (damping, ptof, inverse_product)]
        r_proj, r_adv = mats[0].shape
        self.r_adv, self.r_proj, self.r_fin = r_adv, r_proj, fin.r
        for m in mats[1:]:
            if m.shape != (fin.r, r_proj):
                raise ValueError(f"ptof and inverseProduct must be {fin.r} x {r_proj}, got {m.shape}")
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_plume_create(pre._h, fin._h, r_adv, r_proj, mats[0].ctypes.data,
                                               mats[1].ctypes.data, mats[2].ctypes.data, ctypes.byref(self.params),
                                               ctypes.byref(h)), "cfs_plume_create")
        self._h = h.value
        self.pre, self.fin = pre, fin
        self._keep = ()

    @classmethod
    def from_dense(cls, U_pre: DeviceArray, U_fin: DeviceArray, peeled_dims, damping: np.ndarray, ptof: np.ndarray,
                   inverse_product: np.ndarray, params: Optional[_lib.PlumeParams] = None) -> "PlumeSolver":
        """The uncompressed solver: U_pre, U_fin device FP64 arrays of shape
        (r, 3 Npeel) (mode j = row j, i.e. 3 Npeel x r column-major)."""
        self = cls.__new__(cls)
        X, Y, Z = peeled_dims
        self.full = (X + 2, Y + 2, Z + 2)
        self.r = U_pre.shape[0]
        self.params = params if params is not None else default_params(self.full)
        mats = [np.ascontiguousarray(m, np.float64) for m in (damping, ptof, inverse_product)]
        r_proj, r_adv = mats[0].shape
        self.r_adv, self.r_proj, self.r_fin = r_adv, r_proj, U_fin.shape[0]
        for m in mats[1:]:
            if m.shape != (self.r_fin, r_proj):
                raise ValueError(f"ptof and inverseProduct must be {self.r_fin} x {r_proj}, got {m.shape}")
        n3 = 3 * X * Y * Z
        if U_pre.shape[1] != n3 or U_fin.shape[1] != n3:
            raise ValueError("dense bases must hold 3 Npeel values per mode")
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_plume_create_dense(U_pre.ptr, self.r, U_fin.ptr, self.r_fin, X, Y, Z, r_adv, r_proj,
                                                     mats[0].ctypes.data, mats[1].ctypes.data, mats[2].ctypes.data,
                                                     ctypes.byref(self.params), ctypes.byref(h)),
                   "cfs_plume_create_dense")
        self._h = h.value
        self.pre, self.fin = U_pre, U_fin  # kept alive: the plume reads them every step
        self._keep = ()
        return self

    @classmethod
    def from_operators(cls, pre: CompressedBasis, fin: CompressedBasis, ops, params=None) -> "PlumeSolver":
        """With a ReducedOperators (the reference's .matrix files)."""
        return cls(pre, fin, ops.damping, ops.ptof, ops.inverse_product, params)

    def set_advection(self, mode: int, cubature=None, closure=None):
        _lib.check(_lib.lib().cfs_plume_set_advection(self._h, mode, cubature._h if cubature else None,
                                                      closure.handle if closure else None), "cfs_plume_set_advection")
        self._keep = (cubature, closure)

    def set_exact_after(self, U_adv: DeviceArray):
        """Dense plumes: the exact advection projects onto U_adv (r_adv, 3 Npeel)."""
        _lib.check(_lib.lib().cfs_plume_set_exact_after(self._h, U_adv.ptr), "cfs_plume_set_exact_after")
        self._after = U_adv

    def set_params(self, params: _lib.PlumeParams):
        _lib.check(_lib.lib().cfs_plume_set_params(self._h, ctypes.byref(params)), "cfs_plume_set_params")
        self.params = params

    def advection_state(self):
        """Views of the last step's advection input (q after the projection,
        r_pre values) and output (r_adv values) on the device."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_plume_advection_state(self._h, ctypes.byref(a), ctypes.byref(b)), "advection state")
        return (DeviceArray.view(a.value, (self.r,), np.float64, self),
                DeviceArray.view(b.value, (self.r_adv,), np.float64, self))

    def fields(self):
        """Zeroed device state: velocity (3 per cell), density, heat, q."""
        n = self.full[0] * self.full[1] * self.full[2]
        return (DeviceArray.zeros((3 * n,), np.float32), DeviceArray.zeros((n,), np.float32),
                DeviceArray.zeros((n,), np.float32), DeviceArray.zeros((max(self.r, self.r_fin),), np.float64))

    def add_smoke(self, density: DeviceArray, heat: DeviceArray, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_plume_add_smoke(self._h, density.ptr, heat.ptr, _sptr(stream)), "addSmokeColumn")

    def step(self, velocity: DeviceArray, density: DeviceArray, heat: DeviceArray, q: DeviceArray,
             stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_plume_step(self._h, velocity.ptr, density.ptr, heat.ptr, q.ptr, _sptr(stream)),
                   "stepPlume")

    def frame(self, velocity, density, heat, q, stream: Optional[Stream] = None):
        """One frame of the driver: addSmokeColumn, then stepPlume
        (reducedCompressed3D.cpp:385-386)."""
        self.add_smoke(density, heat, stream)
        self.step(velocity, density, heat, q, stream)

    def set_obstacle(self, U_iop: DeviceArray, U_proj: DeviceArray, iop_T_proj: np.ndarray, ptof_iop: np.ndarray,
                     inverse_iop: np.ndarray):
        """The moving-obstacle matrices: U_iop (r_iop, 3 Npeel) and U_preproject
        (r_proj, 3 Npeel) on the device (kept alive here), U_iop^T U_preproject
        (r_iop x r_proj), and ptof / inverseProduct of the IOP basis (r_fin x r_iop)."""
        r_iop = U_iop.shape[0]
        mats = [np.ascontiguousarray(m, np.float64) for m in (iop_T_proj, ptof_iop, inverse_iop)]
        if U_proj.shape[0] != self.r_proj or U_iop.shape[1] != U_proj.shape[1]:
            raise ValueError("U_preproject must have r_proj modes on the plume's grid")
        if mats[0].shape != (r_iop, self.r_proj):
            raise ValueError(f"U_iop^T U_preproject must be {r_iop} x {self.r_proj}")
        for m in mats[1:]:
            if m.shape != (self.r_fin, r_iop):
                raise ValueError(f"ptof and inverseProduct of the IOP basis must be {self.r_fin} x {r_iop}")
        _lib.check(_lib.lib().cfs_plume_set_obstacle(self._h, r_iop, U_iop.ptr, U_proj.ptr, mats[0].ctypes.data,
                                                     mats[1].ctypes.data, mats[2].ctypes.data),
                   "cfs_plume_set_obstacle")
        self.r_iop = r_iop
        self._obstacle = (U_iop, U_proj)

    def step_obstacle(self, box, velocity: DeviceArray, density: DeviceArray, heat: DeviceArray, q: DeviceArray,
                      stream: Optional[Stream] = None):
        """stepMovingObstacle with the box at its current position."""
        _lib.check(_lib.lib().cfs_plume_step_obstacle(self._h, box.ref, velocity.ptr, density.ptr, heat.ptr, q.ptr,
                                                      _sptr(stream)), "stepMovingObstacle")

    def frame_obstacle(self, box, velocity, density, heat, q, stream: Optional[Stream] = None):
        """addSmokeColumn, stepMovingObstacle, then the box advanced."""
        self.add_smoke(density, heat, stream)
        self.step_obstacle(box, velocity, density, heat, q, stream)
        box.advance()

    def capture(self, velocity, density, heat, q, stream: Stream) -> Graph:
        """A HIP graph of one frame on these arrays."""
        return Graph(stream, lambda: self.frame(velocity, density, heat, q, stream))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cfs_plume_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
