"""The full-rank Stam smoke plume on the device (host side over the C ABI).

Reference (aadjones/CFS_PLUME) -> here:

  FLUID_3D_MIC constructor (FLUID_3D_MIC.cpp:29-160)   default_params, FluidSolver
  addSmokeColumn (FLUID_3D_MIC.cpp:655-662)            FluidSolver.add_smoke
  step (FLUID_3D_MIC.cpp:179-233)                      FluidSolver.step
  appendStreams (FLUID_3D_MIC.cpp:1329-1480)           Snapshots (the six streams, on the device)
  projects/fluid3D runEverytime (fluid3D.cpp:357-385)  FluidSolver.run
  setInitialVelocity (FLUID_3D_MIC.cpp:744-763)        FluidSolver.set_initial_velocity
  stepWithMovingObstacle (FLUID_3D_MIC.cpp:236-310)    FluidSolver.step_obstacle, run_obstacle
                                                       (the IOP stream of appendStreams :1431-1439)

The fields are FP64 (the reference's Real) and stay on the device; a run
fills device snapshot matrices, one column per step, that the basis builder
(`cfs_plume_amd.pipeline`) reads without a host copy.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .basis import _sptr
from .hip import DeviceArray, Stream

STREAMS = ("preadvect", "prediffuse", "preproject", "final_velocity", "pressure", "divergence")
IOP_STREAM = "iop"  # the post-stomp velocity of the obstacle step


def default_params(full_dims, vorticity: float = 0.0) -> _lib.FluidParams:
    """The FLUID_3D_MIC constructor's constants for a full grid, with the
    vorticity the plume drivers set from their cfg ("vorticity", default 0:
    fluid3D.cpp:316,346)."""
    p = _lib.FluidParams()
    X, Y, Z = full_dims
    _lib.lib().cfs_fluid_default_params(X, Y, Z, ctypes.byref(p))
    p.vorticity_eps = float(vorticity)
    return p


@dataclass
class Snapshots:
    """Device snapshot matrices, column-major: column t is step t's peeled
    field (3 values per peeled cell for the velocity streams, 1 for pressure
    and divergence)."""
    arrays: dict
    steps: int
    peeled_cells: int

    def column_ptr(self, name: str, t: int) -> int:
        a = self.arrays[name]
        comps = 1 if name in ("pressure", "divergence") else 3
        return a.ptr + t * comps * self.peeled_cells * 8

    def numpy(self, name: str) -> np.ndarray:
        """[steps][rows] on the host."""
        return self.arrays[name].numpy()


class FluidSolver:
    """FLUID_3D_MIC: full grid X x Y x Z (the reduced bases' peeled dims + 2)."""

    def __init__(self, full_dims, params: Optional[_lib.FluidParams] = None, vorticity: float = 0.0):
        X, Y, Z = (int(d) for d in full_dims)
        self.full = (X, Y, Z)
        self.n = X * Y * Z
        self.peeled_cells = (X - 2) * (Y - 2) * (Z - 2)
        self.params = params if params is not None else default_params(self.full, vorticity)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_fluid_create(X, Y, Z, ctypes.byref(self.params), ctypes.byref(h)),
                   "cfs_fluid_create")
        self._h = h.value
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        _lib.check(_lib.lib().cfs_fluid_fields(self._h, *[ctypes.byref(p) for p in ptrs]), "cfs_fluid_fields")
        self.velocity = DeviceArray.view(ptrs[0].value, (3 * self.n,), np.float64, self)
        self.density = DeviceArray.view(ptrs[1].value, (self.n,), np.float64, self)
        self.heat = DeviceArray.view(ptrs[2].value, (self.n,), np.float64, self)
        self.pressure = DeviceArray.view(ptrs[3].value, (self.n,), np.float64, self)

    def add_smoke(self, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_fluid_add_smoke(self._h, _sptr(stream)), "addSmokeColumn")

    def step(self, snapshots: Optional[Snapshots] = None, t: int = 0, stream: Optional[Stream] = None):
        snap = None
        if snapshots is not None:
            snap = _lib.FluidSnapshots(*[snapshots.column_ptr(name, t) for name in STREAMS])
        _lib.check(_lib.lib().cfs_fluid_step(self._h, ctypes.byref(snap) if snap is not None else None,
                                             _sptr(stream)), "FLUID_3D_MIC::step")

    def set_initial_velocity(self, box, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_fluid_set_initial_velocity(self._h, box.ref, _sptr(stream)), "setInitialVelocity")

    def step_obstacle(self, box, snapshots: Optional[Snapshots] = None, t: int = 0,
                      stream: Optional[Stream] = None):
        """stepWithMovingObstacle with the box at its current position."""
        snap, iop = None, None
        if snapshots is not None:
            snap = _lib.FluidSnapshots(*[snapshots.column_ptr(name, t) for name in STREAMS])
            if IOP_STREAM in snapshots.arrays:
                iop = snapshots.column_ptr(IOP_STREAM, t)
        _lib.check(_lib.lib().cfs_fluid_step_obstacle(self._h, box.ref, ctypes.byref(snap) if snap is not None
                                                      else None, iop, _sptr(stream)),
                   "FLUID_3D_MIC::stepWithMovingObstacle")

    def last_solve(self):
        """(CG iterations, residual maximum) of the last pressure solve."""
        it = ctypes.c_int(0)
        mx = ctypes.c_double(0.0)
        _lib.check(_lib.lib().cfs_fluid_last_solve(self._h, ctypes.byref(it), ctypes.byref(mx)), "last solve")
        return it.value, mx.value

    def new_snapshots(self, steps: int, iop: bool = False) -> Snapshots:
        np_ = self.peeled_cells
        names = STREAMS + ((IOP_STREAM,) if iop else ())
        arrays = {name: DeviceArray.zeros((steps, (1 if name in ("pressure", "divergence") else 3) * np_),
                                          np.float64) for name in names}
        return Snapshots(arrays, steps, np_)

    def run(self, steps: int, stream: Optional[Stream] = None, log=None) -> Snapshots:
        """The full-rank driver (fluid3D.cpp:357-385): per step addSmokeColumn,
        step, and the snapshot streams; returns the device snapshots."""
        snaps = self.new_snapshots(steps)
        for t in range(steps):
            self.add_smoke(stream)
            self.step(snaps, t, stream)
            if log is not None and (t % 10 == 0 or t == steps - 1):
                it, mx = self.last_solve()
                log(f"full-rank step {t + 1}/{steps}: {it} CG iterations, residual max {mx:.2e}")
        return snaps

    def run_obstacle(self, steps: int, box, smoke: bool = True, initial_velocity: bool = False,
                     stream: Optional[Stream] = None, log=None) -> Snapshots:
        """A moving-obstacle run: optionally the box's velocity set into its
        cells first (setInitialVelocity), then per step (passive) smoke, the
        obstacle step with the box where it stands, the seven streams, and the
        box advanced (translate_center, spin, update_time) -- the schedule for
        which buildObstacleMatrices writes its per-step pairs."""
        snaps = self.new_snapshots(steps, iop=True)
        if initial_velocity:
            self.set_initial_velocity(box, stream)
        for t in range(steps):
            if smoke:
                self.add_smoke(stream)
            self.step_obstacle(box, snaps, t, stream)
            box.advance()
            if log is not None and (t % 10 == 0 or t == steps - 1):
                it, mx = self.last_solve()
                log(f"full-rank obstacle step {t + 1}/{steps}: {it} CG iterations, residual max {mx:.2e}")
        return snaps

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cfs_fluid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
