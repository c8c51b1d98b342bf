"""Compressed-versus-uncompressed accuracy of the reduced plume step.

The reference measures what compression costs the reduced simulation by
stepping it once with the dense basis (writing each step's q to qDot.matrix)
and once with the compressed basis, comparing the two q per step
(reducedCompressed3DCompareVelocities.cpp:379-433;
SUBSPACE_FLUID_3D_COMPRESSED_EIGEN::compareSubspace,
SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:271-289: |q_true - q| / |q_true|).

Here both runs step the same reduced plume, q_{t+1} = alpha * U^T A(U q_t) + U^T f
(`ReducedPlume`): the compressed run with the fused decode kernels on the
device, the uncompressed one with the dense FP64 basis the files were
encoded from (the two products on the host, U^T q written into a zero-boundary
FP32 field).  Both use the same device advection A, so the difference is the
compression error (and FP32 rounding of the fields).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .closure import PlumeRun, ReducedPlume, advect_field
from .hip import Stream
from .matrix_io import TrajectoryWriter

__all__ = ["relative_l2_per_step", "DensePlume", "compare_runs", "compare_plume"]


def relative_l2_per_step(q_true: np.ndarray, q: np.ndarray) -> np.ndarray:
    """compareSubspace for every step: |q_true[t] - q[t]| / |q_true[t]|
    (rows [steps][r]; a zero reference row gives inf, or 0 when q matches)."""
    a = np.asarray(q_true, np.float64)
    b = np.asarray(q, np.float64)
    if a.shape != b.shape or a.ndim != 2:
        raise ValueError(f"trajectories of shape {a.shape} and {b.shape}")
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(a, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        out = num / den
    out[(den == 0) & (num == 0)] = 0.0
    return out


class DensePlume:
    """The uncompressed twin of a ReducedPlume: the same grid, advection,
    damping and source, with the dense basis U (FP64 [r][3 X Y Z], row c =
    column c of U, rows interleaved xyz with x fastest)."""

    def __init__(self, plume: ReducedPlume, U_dense: np.ndarray):
        X, Y, Z = plume.U.dims
        self.plume = plume
        self.U = np.ascontiguousarray(U_dense, np.float64)
        if self.U.shape != (plume.U.r, 3 * X * Y * Z):
            raise ValueError(f"dense basis of shape {self.U.shape} for r = {plume.U.r}, dims {plume.U.dims}")
        self._full = np.zeros((Z + 2, Y + 2, X + 2, 3), np.float32)

    def _peel(self, full: np.ndarray) -> np.ndarray:
        X, Y, Z = self.plume.U.dims
        return full.reshape(Z + 2, Y + 2, X + 2, 3)[1:-1, 1:-1, 1:-1].reshape(-1)

    def forcing(self, run: PlumeRun) -> np.ndarray:
        """b = U^T f of the run's source (FP64 [r])."""
        return self.U @ self._peel(self.plume.forcing_field(run)).astype(np.float64)

    def step(self, q: np.ndarray, b: np.ndarray, stream: Optional[Stream] = None) -> np.ndarray:
        """alpha * U^T A(U q) + b, with A on the device."""
        p = self.plume
        X, Y, Z = p.U.dims
        self._full[1:-1, 1:-1, 1:-1] = (q @ self.U).astype(np.float32).reshape(Z, Y, X, 3)
        p.fa.upload(self._full.reshape(-1))
        advect_field(p.fa, p.fb, p.full, p.dt0, stream)
        if stream is not None:
            stream.synchronize()
        adv = self._peel(p.fb.numpy()).astype(np.float64)
        return p.alpha * (self.U @ adv) + b

    def rollout(self, run: PlumeRun, steps: int, stream: Optional[Stream] = None) -> np.ndarray:
        """q_t for t = 0..steps (q_0 = 0): FP64 [steps+1][r]."""
        b = self.forcing(run)
        out = np.zeros((steps + 1, self.U.shape[0]))
        for t in range(steps):
            out[t + 1] = self.step(out[t], b, stream)
        return out


def compare_runs(plume: ReducedPlume, U_dense: np.ndarray, runs, steps: int, out_prefix: Optional[str] = None):
    """Per-step relative L2 error of the compressed run's q against the
    uncompressed run's, for each run: FP64 [runs][steps] (steps 1..steps).
    With out_prefix, each run's trajectories are written as
    <prefix>.<i>.qDot.matrix (uncompressed) and <prefix>.<i>.compressed.qDot.matrix."""
    dense = DensePlume(plume, U_dense)
    b_dev = plume.forcing(runs)
    q_c = plume.rollout_exact(b_dev, len(runs), steps)
    errs = np.zeros((len(runs), steps))
    for i, run in enumerate(runs):
        q_u = dense.rollout(run, steps)
        errs[i] = relative_l2_per_step(q_u[1:], q_c[i, 1:])
        if out_prefix:
            for name, traj in ((f"{out_prefix}.{i}.qDot.matrix", q_u), (f"{out_prefix}.{i}.compressed.qDot.matrix", q_c[i])):
                w = TrajectoryWriter(name, steps, traj.shape[1])
                for t in range(1, steps + 1):
                    w.append(traj[t])
    return errs


def compare_plume(data, steps: int, stream: Optional[Stream] = None, out_prefix: Optional[str] = None,
                  exact: bool = True) -> dict:
    """The reference's measurement on a real plume (reducedCompressed3DCompareVelocities.cpp:
    379-433): the whole reduced step (`stepPlume`, addSmokeColumn every frame,
    the cubature advection) run with the compressed U_preadvect / U_final on
    the device and with the dense bases they were encoded from (the
    uncompressed solver, dense products on the device), from rest;
    compareSubspace per step, and the reduced velocity of each run against
    the full-rank run's final velocity of the same step (diffGroundTruth).

    `data` is a pipeline.PlumeData.  With a moving obstacle (data.box), every
    run steps with stepMovingObstacle from the box's starting position, the
    IOP products of `data.operators` and U_iop / U_preproject dense on the
    device.  Returns per-step arrays."""
    from .cubature import Cubature
    from .hip import DeviceArray
    from .pipeline import compressed_bases
    from .plume import ADVECT_CUBATURE, ADVECT_EXACT, PlumeSolver, default_params

    dims = data.peeled
    ops = data.operators
    D, P, I = ops["damping.peeled"], ops["projected.ptof"], ops["inverseProduct"]
    box0 = getattr(data, "box", None)
    if box0 is not None:
        # the plain projection matrices are not used by the obstacle step
        P_iop, I_iop = P, I
        P = I = np.zeros((data.bases["final"].r, D.shape[0]))
    prm = default_params(data.full)
    prm.vorticity_eps = 0.0  # the drivers' cfg value ("vorticity", default 0)
    cub = Cubature.from_rows(data.cubature.cells, data.cubature.weights, dims, data.bases["prediffuse"].numpy())
    comp = compressed_bases(data)
    runs = {
        "compressed": PlumeSolver(comp["preadvect"], comp["final"], D, P, I, prm),
        "uncompressed": PlumeSolver.from_dense(data.bases["preadvect"].U, data.bases["final"].U, dims, D, P, I, prm),
    }
    if exact:
        # the uncompressed solver with the exact reduced advection U_prediffuse^T A(U_preadvect q):
        # what the cubature approximates
        runs["uncompressed_exact"] = PlumeSolver.from_dense(data.bases["preadvect"].U, data.bases["final"].U, dims,
                                                            D, P, I, prm)
        runs["uncompressed_exact"].set_exact_after(data.bases["prediffuse"].U)
    truth = data.snapshots.arrays["final_velocity"].numpy()
    r_fin = data.bases["final"].r
    out = {}
    for name, solver in runs.items():
        if name == "uncompressed_exact":
            solver.set_advection(ADVECT_EXACT)
        else:
            solver.set_advection(ADVECT_CUBATURE, cubature=cub)
        box = None
        if box0 is not None:
            solver.set_obstacle(data.bases["iop"].U, data.bases["preproject"].U, ops["U'U.iop.subspace"], P_iop,
                                I_iop)
            box = box0.copy()
        vel, dens, heat, q = solver.fields()
        qs = np.zeros((steps, r_fin))
        vdiff = np.zeros(steps)
        X, Y, Z = data.full
        for t in range(steps):
            if box is not None:
                solver.frame_obstacle(box, vel, dens, heat, q, stream)
            else:
                solver.frame(vel, dens, heat, q, stream)
            qs[t] = q.numpy()[:r_fin]
            v = vel.numpy().reshape(Z, Y, X, 3)[1:-1, 1:-1, 1:-1].reshape(-1).astype(np.float64)
            if t < truth.shape[0]:
                vdiff[t] = np.linalg.norm(v - truth[t]) / max(np.linalg.norm(truth[t]), 1e-300)
        out[name] = {"q": qs, "velocity_vs_full_rank": vdiff}
        solver.close()
    if out_prefix:
        for name in runs:
            w = TrajectoryWriter(f"{out_prefix}.{name}.qDot.matrix", steps, r_fin)
            for row in out[name]["q"]:
                w.append(row)
    out["q_rel_l2"] = relative_l2_per_step(out["uncompressed"]["q"], out["compressed"]["q"])
    cub.close()
    return out
