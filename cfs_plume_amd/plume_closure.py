"""The learned closure on the real reduced plume (config C5 on the plume's
own bases): an MLP replaces the cubature advection inside `stepPlume`,
trained on the pairs the reduced solver itself produces.

The reduced solver is the device `stepPlume` (plume.PlumeSolver) on the
compressed U_preadvect / U_final of a plume built by pipeline.build_plume_data,
with the staged cubature as its advection (reducedAdvectCompressionFriendly:
U_preadvect coordinates in, U_prediffuse coordinates out).  Runs are seeded
as the C5 configuration asks: the smoke column shifted by offsets uniform in
+-8 cells in x and z, buoyancy scaled uniformly in [0.5, 2].  Every frame's
advection input q_pre (zero-padded to r_adv) and output q_adv are the
training pairs (q_pre, q_adv - q_pre); the same pairs at states perturbed
around them are evaluated with the cubature on the device.  The trained
closure (closure.Closure with r = r_adv) then runs inside the step in place
of the cubature, and held-out runs are compared frame by frame with the
cubature runs -- the cubature is the closure's judge and fallback.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .closure import Closure, timed
from .cubature import Cubature
from .hip import DeviceArray, Stream
from .pipeline import PlumeData, compressed_bases
from .plume import ADVECT_CLOSURE, ADVECT_CUBATURE, PlumeSolver, default_params


@dataclass
class PlumeRunSpec:
    offset_x: float
    offset_z: float
    buoyancy_scale: float


def seeded_plume_runs(n: int, seed: int):
    rng = np.random.default_rng(seed)
    return [PlumeRunSpec(float(rng.uniform(-8, 8)), float(rng.uniform(-8, 8)), float(rng.uniform(0.5, 2.0)))
            for _ in range(n)]


class PlumeClosureExperiment:
    def __init__(self, data: PlumeData, frames: int = 100, stream: Optional[Stream] = None):
        self.data = data
        self.frames = frames
        self.stream = stream or Stream()
        dims = data.peeled
        ops = data.operators
        self.bases = compressed_bases(data)
        self.cub = Cubature.from_rows(data.cubature.cells, data.cubature.weights, dims,
                                      data.bases["prediffuse"].numpy())
        self.base_params = default_params(data.full)
        self.base_params.vorticity_eps = 0.0
        self.solver = PlumeSolver(self.bases["preadvect"], self.bases["final"], ops["damping.peeled"],
                                  ops["projected.ptof"], ops["inverseProduct"], self.base_params)
        self.solver.set_advection(ADVECT_CUBATURE, cubature=self.cub)
        self.r_pre, self.r_adv, self.r_fin = self.solver.r, self.solver.r_adv, self.solver.r_fin
        self.dt0 = self.base_params.dt / self.base_params.dx
        self.scratch = DeviceArray.zeros((3 * int(np.prod(data.full)),), np.float32)

    def _params(self, spec: PlumeRunSpec):
        p = default_params(self.data.full)
        p.vorticity_eps = 0.0
        p.buoyancy = self.base_params.buoyancy * spec.buoyancy_scale
        p.smoke_offset_x, p.smoke_offset_z = spec.offset_x, spec.offset_z
        return p

    def run(self, spec: PlumeRunSpec, closure: Optional[Closure] = None, collect: bool = False,
            check_every: int = 0, tol: float = 0.01):
        """One run from rest: q in final coordinates per frame ([frames][r_fin]),
        and with collect the advection pairs ([frames][r_adv] each).

        With a closure and check_every = k > 0, every k-th frame is stepped
        with the cubature (the judge) and the closure is evaluated at the same
        advection input; when they differ by more than `tol` (relative L2)
        the cubature stays the advection until the next check (the fallback).
        Returns (finals, fallback frames) in that case."""
        s = self.solver
        use_closure = closure is not None
        s.set_params(self._params(spec))
        vel, dens, heat, q = s.fields()
        qa, qb = s.advection_state()
        finals = np.zeros((self.frames, self.r_fin))
        xs = np.zeros((self.frames, self.r_adv)) if collect else None
        ys = np.zeros((self.frames, self.r_adv)) if collect else None
        fallback = 0
        fallback_on = False
        probe = DeviceArray.zeros((self.r_adv,), np.float64) if check_every else None
        for t in range(self.frames):
            checked = use_closure and check_every > 0 and t % check_every == 0
            by_cubature = (not use_closure) or checked or fallback_on
            if by_cubature:
                s.set_advection(ADVECT_CUBATURE, cubature=self.cub)
            else:
                s.set_advection(ADVECT_CLOSURE, closure=closure)
            s.frame(vel, dens, heat, q, self.stream)
            finals[t] = q.numpy()[:self.r_fin]
            if checked:
                # the closure at this frame's advection input (zero-padded on the device)
                closure.step(qa, probe, 1, 1.0, None, self.stream)
                ref = qb.numpy()
                fallback_on = bool(np.linalg.norm(probe.numpy() - ref) > tol * np.linalg.norm(ref))
            elif use_closure and by_cubature:
                fallback += 1
            if collect:
                xs[t, :self.r_pre] = qa.numpy()
                ys[t] = qb.numpy()
        if collect:
            return finals, xs, ys
        return (finals, fallback) if (use_closure and check_every > 0) else finals

    def cubature_at(self, states: np.ndarray) -> np.ndarray:
        """The cubature advection at given U_preadvect states ([N][r_pre])."""
        n = states.shape[0]
        qin = DeviceArray.from_numpy(np.ascontiguousarray(states[:, :self.r_pre]))
        qout = DeviceArray.zeros((n, self.r_adv), np.float64)
        for i in range(n):
            self.cub.advect(self.bases["preadvect"], (qin, i), (qout, i), self.scratch, self.dt0,
                            stream=self.stream)
        return qout.numpy()

    def train(self, runs, epochs: int = 600, noise: float = 0.2, seed: int = 5, log=None):
        t0 = time.perf_counter()
        X, Y = [], []
        for spec in runs:
            _, xs, ys = self.run(spec, collect=True)
            X.append(xs)
            Y.append(ys)
        X, Y = np.concatenate(X), np.concatenate(Y)
        if noise > 0:
            rng = np.random.default_rng(11)
            scale = np.linalg.norm(X, axis=1, keepdims=True) / np.sqrt(self.r_pre)
            Xp = X.copy()
            Xp[:, :self.r_pre] += noise * rng.normal(size=(X.shape[0], self.r_pre)) * scale
            Yp = self.cubature_at(Xp)
            X, Y = np.concatenate([X, Xp]), np.concatenate([Y, Yp])
        t_data = time.perf_counter() - t0
        cl = Closure(self.r_adv, max_batch=1024, seed=seed)
        t0 = time.perf_counter()
        losses = cl.fit(X, Y - X, epochs=epochs, batch=512, lr=2e-3, lr_final=1e-5, seed=7,
                        log=(lambda ep, m: log(f"plume closure epoch {ep}: mse {m:.3e}") if ep % 50 == 0 else None)
                        if log else None)
        return cl, {"samples": int(X.shape[0]), "losses": losses, "data_s": t_data,
                    "train_s": time.perf_counter() - t0}

    def evaluate(self, closure: Closure, runs) -> dict:
        """Per held-out run, q (final coordinates) of the closure run against
        the cubature run, relative L2 per frame; the same with the closure
        checked against the cubature every 10 frames (fallback above 1 %);
        and the closure's one-step error along the cubature trajectory."""
        drift = np.zeros((len(runs), self.frames))
        checked = np.zeros((len(runs), self.frames))
        one_step = []
        fallbacks = 0
        for k, spec in enumerate(runs):
            ref, xs, ys = self.run(spec, collect=True)
            got = self.run(spec, closure)
            drift[k] = np.linalg.norm(got - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
            chk, fb = self.run(spec, closure, check_every=10)
            fallbacks += fb
            checked[k] = np.linalg.norm(chk - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
            # the closure's one-step error on the cubature trajectory (teacher-forced)
            xin = DeviceArray.from_numpy(xs)
            pred = DeviceArray.zeros(xs.shape, np.float64)
            closure.step(xin, pred, xs.shape[0], 1.0, None, self.stream)
            p_ = pred.numpy()
            one_step.append(np.linalg.norm(p_ - ys, axis=1) / np.maximum(np.linalg.norm(ys, axis=1), 1e-300))
        return {"drift": drift, "checked_drift": checked, "fallback_frames": fallbacks,
                "one_step": np.concatenate(one_step)}

    def step_times(self, closure: Closure, reps: int = 200) -> dict:
        """The advection stage alone (n = 1, HIP events): cubature
        (unproject + key cells + sum) against the closure's fused step."""
        qin = DeviceArray.from_numpy(np.zeros(self.r_adv))
        qin.upload(np.random.default_rng(1).standard_normal(self.r_adv) * 0.01)
        qout = DeviceArray.zeros((self.r_adv,), np.float64)
        pre = self.bases["preadvect"]
        ms_cub = timed(lambda s: self.cub.advect(pre, qin, qout, self.scratch, self.dt0, stream=s), reps, self.stream)
        ms_cl = timed(lambda s: closure.step(qin, qout, 1, 1.0, None, s), reps, self.stream)
        return {"cubature_ms": ms_cub, "closure_ms": ms_cl}

    def close(self):
        self.solver.close()
        self.cub.close()
