"""Reduced advection and its learned closure (host side over the C ABI).

Reference (aadjones/CFS_PLUME) -> here:

  reducedAdvectCompressionFriendly (cubature,        reduced_advect (exact: every cell is a
    SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1781-1854)   key cell, weights 1)
  VECTOR3_FIELD_3D::advect (VECTOR3_FIELD_3D.cpp:    advect_field
    2001-2077)
  stepPlume's advection stage (:481)                 ReducedPlume.rollout_exact / rollout_closure

The closure (no counterpart in the reference) is an MLP r -> 256 -> 256 -> 256
-> r with GELU, BF16 weights and FP32 accumulation, trained on the GPU by the
library's own forward/backward kernels and Adam (Closure.train), and used in
place of the advection with the exact step as fallback
(ReducedPlume.rollout_closure_checked).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .basis import CompressedBasis, _sptr
from .hip import DeviceArray, Event, Stream

HIDDEN = 256


def advect_field(vel: DeviceArray, out: DeviceArray, full_dims, dt0: float, stream: Optional[Stream] = None):
    """out = VECTOR3_FIELD_3D::advect of the full-grid field vel by itself."""
    X, Y, Z = full_dims
    _lib.check(_lib.lib().cfs_advect_field(vel.ptr, out.ptr, X, Y, Z, float(dt0), _sptr(stream)), "cfs_advect_field")
    return out


def advect_scalar(vel: DeviceArray, old: DeviceArray, out: DeviceArray, full_dims, dt0: float,
                  stream: Optional[Stream] = None):
    """out = VECTOR3_FIELD_3D::advect of the full-grid scalar field `old` along
    vel (the density and heat advection of advectHeatAndDensityStam)."""
    X, Y, Z = full_dims
    _lib.check(_lib.lib().cfs_advect_scalar(vel.ptr, old.ptr, out.ptr, X, Y, Z, float(dt0), _sptr(stream)),
               "cfs_advect_scalar")
    return out


def _ptr(a: Optional[DeviceArray], offset_elems: int = 0):
    if a is None:
        return None
    return a.ptr + offset_elems * a.dtype.itemsize


def reduced_advect(U: CompressedBasis, q_in, q_out, field_a: DeviceArray, field_b: DeviceArray, dt0: float,
                   alpha: float = 1.0, forcing=None, stream: Optional[Stream] = None):
    """q_out = alpha * U^T A(U q_in) + forcing on the device.

    q_in / q_out / forcing: DeviceArray or (DeviceArray, row) pairs addressing
    row `row` of an [n][r] array."""
    def p(a):
        if a is None:
            return None
        if isinstance(a, tuple):
            return _ptr(a[0], a[1] * U.r)
        return a.ptr
    _lib.check(_lib.lib().cfs_reduced_advect(U._h, p(q_in), p(q_out), field_a.ptr, field_b.ptr, float(dt0),
                                             float(alpha), p(forcing), _sptr(stream)), "cfs_reduced_advect")


class Closure:
    """The learned advection closure (device-resident parameters)."""

    def __init__(self, r: int, max_batch: int = 1024, seed: int = 1):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_closure_create(r, HIDDEN, max_batch, seed, ctypes.byref(h)), "cfs_closure_create")
        self.handle = h.value
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().cfs_closure_get_dims(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                   ctypes.byref(n)), "cfs_closure_get_dims")
        self.r, self.rp, self.hidden, self.n_params = a.value, b.value, c.value, n.value
        self.max_batch = max_batch
        self._loss = DeviceArray.zeros((1,), np.float64)
        self.L = None

    # -- parameters ---------------------------------------------------------
    def layer_shapes(self):
        K = [self.rp, HIDDEN, HIDDEN, HIDDEN]
        M = [HIDDEN, HIDDEN, HIDDEN, self.rp]
        return list(zip(K, M))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        _lib.check(_lib.lib().cfs_closure_get_params(self.handle, out.ctypes.data), "cfs_closure_get_params")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _lib.check(_lib.lib().cfs_closure_set_params(self.handle, p.ctypes.data), "cfs_closure_set_params")

    def set_scales(self, s_in: np.ndarray, s_out: np.ndarray):
        a = np.ascontiguousarray(s_in, np.float32)
        b = np.ascontiguousarray(s_out, np.float32)
        assert a.size == self.r and b.size == self.r
        _lib.check(_lib.lib().cfs_closure_set_scales(self.handle, a.ctypes.data, b.ctypes.data), "cfs_closure_set_scales")

    def set_linear(self, L: Optional[np.ndarray]):
        """FP32 linear skip (r x r, input-major) added to the network output."""
        if L is None:
            _lib.check(_lib.lib().cfs_closure_set_linear(self.handle, None), "cfs_closure_set_linear")
            self.L = None
            return
        a = np.ascontiguousarray(L, np.float32)
        assert a.shape == (self.r, self.r)
        _lib.check(_lib.lib().cfs_closure_set_linear(self.handle, a.ctypes.data), "cfs_closure_set_linear")
        self.L = a

    def fit(self, q: np.ndarray, d: np.ndarray, epochs: int, batch: int = 512, lr: float = 2e-3,
            lr_final: float = 1e-5, seed: int = 0, linear: bool = True, log=None):
        """Fits the closure to advection increments d at states q ([N][r]):
        per-component scales, the least-squares linear skip (FP64 on the host)
        and the network on the residual (Adam on the device).  Returns the
        per-epoch mean squared error of the scaled residual."""
        s_in, s_out = fit_scales(q, d)
        self.set_scales(s_in, s_out)
        x = q * s_in
        t = d / s_out
        if linear:
            L = np.linalg.lstsq(x, t, rcond=None)[0]
            self.set_linear(L)
            t = t - x @ L.astype(np.float32).astype(np.float64)
        else:
            self.set_linear(None)
        return self.train(x, t, epochs, batch, lr, lr_final, seed, log)

    def get_scales(self):
        a = np.empty(self.r, np.float32)
        b = np.empty(self.r, np.float32)
        _lib.check(_lib.lib().cfs_closure_get_scales(self.handle, a.ctypes.data, b.ctypes.data), "cfs_closure_get_scales")
        return a, b

    # -- network ------------------------------------------------------------
    def pad(self, a: np.ndarray) -> np.ndarray:
        """[n][r] -> [n][r_padded] FP32 with zero padding columns."""
        a = np.atleast_2d(a)
        out = np.zeros((a.shape[0], self.rp), np.float32)
        out[:, :self.r] = a
        return out

    def forward(self, x: DeviceArray, n: int, y: Optional[DeviceArray] = None,
                stream: Optional[Stream] = None) -> DeviceArray:
        if y is None:
            y = DeviceArray((n, self.rp), np.float32)
        _lib.check(_lib.lib().cfs_closure_forward(self.handle, x.ptr, y.ptr, n, _sptr(stream)), "cfs_closure_forward")
        return y

    def gradients(self, x: DeviceArray, t: DeviceArray, n: int, stream: Optional[Stream] = None):
        g = DeviceArray((self.n_params,), np.float32)
        _lib.check(_lib.lib().cfs_closure_gradients(self.handle, x.ptr, t.ptr, n, g.ptr, self._loss.ptr, _sptr(stream)),
                   "cfs_closure_gradients")
        return g.numpy(), float(self._loss.numpy()[0])

    def train_step(self, x: DeviceArray, t: DeviceArray, rows: DeviceArray, n: int, lr: float,
                   loss: Optional[DeviceArray] = None, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_closure_train_step(self.handle, x.ptr, t.ptr, rows.ptr, n, float(lr),
                                                     loss.ptr if loss is not None else None, _sptr(stream)),
                   "cfs_closure_train_step")

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 512, lr: float = 1e-3,
              lr_final: float = 1e-5, seed: int = 0, log=None):
        """Adam on the scaled pairs (x, t) ([N][r] each, host arrays); the
        learning rate decays geometrically from lr to lr_final.  Returns the
        mean squared error of every epoch."""
        N = x.shape[0]
        batch = min(batch, self.max_batch, N)
        dx = DeviceArray.from_numpy(self.pad(x))
        dt = DeviceArray.from_numpy(self.pad(t))
        rng = np.random.default_rng(seed)
        steps_per_epoch = N // batch
        total = max(1, epochs * steps_per_epoch)
        gamma = (lr_final / lr) ** (1.0 / total)
        losses = DeviceArray.zeros((max(1, steps_per_epoch),), np.float64)
        out, k = [], 0
        for ep in range(epochs):
            perm = rng.permutation(N).astype(np.int32)
            rows = DeviceArray.from_numpy(perm[:steps_per_epoch * batch])
            for s in range(steps_per_epoch):
                self.train_step(dx, dt, _Rows(rows, s * batch), batch, lr * gamma ** k, _Row(losses, s))
                k += 1
            mse = float(losses.numpy().sum()) / (steps_per_epoch * batch * self.r)
            out.append(mse)
            if log:
                log(ep, mse)
        return out

    def step(self, q_in, q_out, n: int = 1, alpha: float = 1.0, forcing=None, stream: Optional[Stream] = None):
        """q_out = alpha * (q_in + s_out * MLP(s_in * q_in)) + forcing (FP64 [n][r])."""
        _lib.check(_lib.lib().cfs_closure_step(self.handle, _rowptr(q_in, self.r), _rowptr(q_out, self.r), n,
                                               float(alpha), _rowptr(forcing, self.r), _sptr(stream)),
                   "cfs_closure_step")

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().cfs_closure_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Rows:
    """Device pointer to element `offset` of an int32 array."""

    def __init__(self, a: DeviceArray, offset: int):
        self.ptr = a.ptr + offset * 4


class _Row:
    def __init__(self, a: DeviceArray, offset: int):
        self.ptr = a.ptr + offset * a.dtype.itemsize


def _rowptr(a, r):
    if a is None:
        return None
    if isinstance(a, tuple):
        return a[0].ptr + a[1] * r * a[0].dtype.itemsize
    return a.ptr


@dataclass
class PlumeRun:
    """A seeded reduced-plume run: a buoyant source of radius 4 cells at the
    bottom centre, offset in x and z, with a buoyancy factor (the plume
    source of the reference's configs, FLUID_3D_MIC.cpp:674-703)."""
    offset_x: float
    offset_z: float
    buoyancy: float


class ReducedPlume:
    """Reduced-space plume stand-in on one compressed basis.

    q_{t+1} = alpha * Adv(q_t) + b_run, with Adv the exact reduced advection
    (U^T A(U q)) or the learned closure, alpha a uniform damping standing in
    for the reduced diffusion, and b_run = U^T f the projected buoyancy of
    the run's source (upward force `force * buoyancy` inside the source
    sphere)."""

    def __init__(self, U: CompressedBasis, dt0: float = 1.0, alpha: float = 0.98, force: float = 0.05,
                 radius: float = 4.0, cubature=None):
        self.U, self.dt0, self.alpha, self.force, self.radius = U, dt0, alpha, force, radius
        X, Y, Z = U.dims
        self.full = (X + 2, Y + 2, Z + 2)
        n = 3 * self.full[0] * self.full[1] * self.full[2]
        self.fa = DeviceArray((n,), np.float32)
        self.fb = DeviceArray((n,), np.float32)
        # the plain reduced solver's advection: the cubature when one is set
        # (reducedAdvectCompressionFriendly), else the exact U^T A(U q)
        self.cubature = cubature

    def advect(self, q_in, q_out, alpha: float, forcing=None, stream: Optional[Stream] = None):
        """q_out = alpha * Adv(q_in) + forcing with the plain solver's advection."""
        if self.cubature is not None:
            self.cubature.advect(self.U, q_in, q_out, self.fa, self.dt0, alpha, forcing, stream)
        else:
            reduced_advect(self.U, q_in, q_out, self.fa, self.fb, self.dt0, alpha, forcing, stream)

    def rollout(self, b: DeviceArray, runs: int, steps: int, stream: Optional[Stream] = None) -> np.ndarray:
        """Like rollout_exact with the plain solver's advection (the cubature
        when one is set)."""
        r = self.U.r
        traj = DeviceArray.zeros((runs * (steps + 1), r), np.float64)
        for i in range(runs):
            for t in range(steps):
                self.advect((traj, i * (steps + 1) + t), (traj, i * (steps + 1) + t + 1), self.alpha, (b, i), stream)
        return traj.numpy().reshape(runs, steps + 1, r)

    def increments_at(self, states: np.ndarray, stream: Optional[Stream] = None) -> np.ndarray:
        """Adv(q) - q for every row of `states` with the plain solver's advection."""
        r = self.U.r
        n = states.shape[0]
        qin = DeviceArray.from_numpy(np.ascontiguousarray(states, np.float64))
        qout = DeviceArray((n, r), np.float64)
        for i in range(n):
            self.advect((qin, i), (qout, i), 1.0, None, stream)
        return qout.numpy() - states

    def make_cubature(self, states: np.ndarray, candidates_per_pass: int = 5000, passes: int = 6,
                      tolerance: float = 0.01, log=None):
        """Fits a cubature (generateImportanceSampledCubature) to the exact
        reduced advection at `states` ([S][r], the training samples) and sets
        it as this plume's advection.  Returns the generator's result."""
        from . import cubature as CB
        rows = CB.dense_rows(self.U)
        post = self.exact_increments(states) + states
        res = CB.generate_cubature(rows, rows, self.U.dims, states, post, self.dt0,
                                   candidates_per_pass=candidates_per_pass, passes=passes, tolerance=tolerance,
                                   log=log)
        self.cubature = CB.Cubature.from_cells(res.cells, res.weights, self.U)
        return res

    @staticmethod
    def seeded_runs(n: int, seed: int) -> Sequence[PlumeRun]:
        rng = np.random.default_rng(seed)
        return [PlumeRun(float(rng.uniform(-8, 8)), float(rng.uniform(-8, 8)), float(rng.uniform(0.5, 2.0)))
                for _ in range(n)]

    def forcing_field(self, run: PlumeRun) -> np.ndarray:
        """Full-grid FP32 force field of the run (interior cells only)."""
        FX, FY, FZ = self.full
        z, y, x = np.meshgrid(np.arange(FZ), np.arange(FY), np.arange(FX), indexing="ij")
        cx, cy, cz = FX / 2.0 + run.offset_x, 1.0 + self.radius + 1.0, FZ / 2.0 + run.offset_z
        inside = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= self.radius ** 2
        inside &= (x > 0) & (x < FX - 1) & (y > 0) & (y < FY - 1) & (z > 0) & (z < FZ - 1)
        f = np.zeros((FZ, FY, FX, 3), np.float32)
        f[..., 1] = np.where(inside, self.force * run.buoyancy, 0.0)
        return f.reshape(-1)

    def forcing(self, runs: Sequence[PlumeRun]) -> DeviceArray:
        """b for every run: FP64 [n][r] on the device (projected on the GPU)."""
        out = DeviceArray((len(runs), self.U.r), np.float64)
        q = DeviceArray((self.U.r,), np.float64)
        fa = self.fa
        for i, run in enumerate(runs):
            fa.upload(self.forcing_field(run))
            self.U.project(fa, q)
            _copy_d2d(out, i * self.U.r, q, 0, self.U.r)
        return out

    def rollout_exact(self, b: DeviceArray, runs: int, steps: int, stream: Optional[Stream] = None) -> np.ndarray:
        """q_t for t = 0..steps of every run (q_0 = 0): host FP64 [runs][steps+1][r]."""
        r = self.U.r
        traj = DeviceArray.zeros((runs * (steps + 1), r), np.float64)
        for i in range(runs):
            for t in range(steps):
                reduced_advect(self.U, (traj, i * (steps + 1) + t), (traj, i * (steps + 1) + t + 1), self.fa, self.fb,
                               self.dt0, self.alpha, (b, i), stream)
        return traj.numpy().reshape(runs, steps + 1, r)

    def exact_increments(self, states: np.ndarray, stream: Optional[Stream] = None) -> np.ndarray:
        """Adv(q) - q for every row of `states` ([N][r], host FP64), with the
        exact reduced advection on the device."""
        r = self.U.r
        n = states.shape[0]
        qin = DeviceArray.from_numpy(np.ascontiguousarray(states, np.float64))
        qout = DeviceArray((n, r), np.float64)
        for i in range(n):
            reduced_advect(self.U, (qin, i), (qout, i), self.fa, self.fb, self.dt0, 1.0, None, stream)
        return qout.numpy() - states

    def rollout_closure(self, closure: Closure, b: DeviceArray, runs: int, steps: int,
                        stream: Optional[Stream] = None) -> np.ndarray:
        """All runs advanced together by the closure step (n = runs per call)."""
        r = self.U.r
        qs = [DeviceArray.zeros((runs, r), np.float64) for _ in range(2)]
        out = np.zeros((runs, steps + 1, r))
        for t in range(steps):
            closure.step(qs[t & 1], qs[(t + 1) & 1], runs, self.alpha, b, stream)
            out[:, t + 1] = qs[(t + 1) & 1].numpy()
        return out

    def rollout_closure_checked(self, closure: Closure, b: DeviceArray, run: int, steps: int, check_every: int,
                                tol: float = 0.01, stream: Optional[Stream] = None):
        """One run with the closure; every `check_every` steps the plain
        solver's step (the cubature when one is set) is also taken from the
        same state and replaces the closure's result when their relative
        difference exceeds `tol` (the fallback).  Returns (trajectory
        [steps+1][r], number of fallbacks)."""
        r = self.U.r
        q = DeviceArray.zeros((r,), np.float64)
        qc = DeviceArray.zeros((r,), np.float64)
        qe = DeviceArray.zeros((r,), np.float64)
        out = np.zeros((steps + 1, r))
        falls = 0
        for t in range(steps):
            closure.step(q, qc, 1, self.alpha, (b, run), stream)
            if check_every and (t + 1) % check_every == 0:
                self.advect(q, qe, self.alpha, (b, run), stream)
                a, e = qc.numpy(), qe.numpy()
                if np.linalg.norm(a - e) > tol * max(np.linalg.norm(e), 1e-300):
                    falls += 1
                    _copy_d2d(qc, 0, qe, 0, r)
            q, qc = qc, q
            out[t + 1] = q.numpy()
        return out, falls


def _copy_d2d(dst: DeviceArray, dst_off: int, src: DeviceArray, src_off: int, n: int):
    from .hip import hip, check, hipMemcpyDeviceToDevice
    isz = dst.dtype.itemsize
    check(hip().hipMemcpy(dst.ptr + dst_off * isz, src.ptr + src_off * isz, n * isz, hipMemcpyDeviceToDevice),
          "hipMemcpy D2D")


def increments(traj: np.ndarray, b: np.ndarray, alpha: float):
    """Training pairs from exact trajectories: q_{t+1} = alpha (q_t + d_t) + b
    gives the advection increment d_t = (q_{t+1} - b) / alpha - q_t."""
    q = traj[:, :-1]
    d = (traj[:, 1:] - b[:, None, :]) / alpha - q
    return q.reshape(-1, q.shape[-1]), d.reshape(-1, d.shape[-1])


def sample_states(traj: np.ndarray, n: int) -> np.ndarray:
    """n states spread evenly over the steps t >= 1 of all runs of `traj`
    ([runs][steps+1][r]): training samples for a cubature fit."""
    q = traj[:, 1:].reshape(-1, traj.shape[-1])
    idx = np.linspace(0, len(q) - 1, n).round().astype(np.int64)
    return np.ascontiguousarray(q[idx])


def cubature_training_states(plume: "ReducedPlume", b: DeviceArray, runs: int, steps: int, samples: int,
                             noise: float = 0.1, seed: int = 3) -> np.ndarray:
    """States a plume's cubature is fitted to: samples / 2 states spread over
    exact rollouts of `runs` runs, and the same states perturbed by `noise`
    times |q| / sqrt(r)."""
    traj = plume.rollout_exact(b, runs, steps)
    st = sample_states(traj, samples // 2)
    rng = np.random.default_rng(seed)
    pert = st + noise * rng.normal(size=st.shape) * np.linalg.norm(st, axis=1, keepdims=True) / np.sqrt(st.shape[1])
    return np.concatenate([st, pert])


def train_closure(plume: "ReducedPlume", b: DeviceArray, runs: int, steps: int, epochs: int, noise: float = 0.1,
                  seed: int = 5, max_batch: int = 1024, log=None):
    """Trains a closure on trajectories of the plain reduced solver (the
    plume's advection: the cubature when one is set): `runs` x `steps` pairs
    (q_t, Adv(q_t) - q_t), plus the same at states perturbed by `noise` times
    |q| / sqrt(r) around them.  Returns (closure, {samples, losses, seconds})."""
    import time
    r = plume.U.r
    t0 = time.perf_counter()
    traj = plume.rollout(b, runs, steps)
    q, d = increments(traj, b.numpy(), plume.alpha)
    if noise > 0:
        rng = np.random.default_rng(11)
        qp = q + noise * rng.normal(size=q.shape) * np.linalg.norm(q, axis=1, keepdims=True) / np.sqrt(r)
        q, d = np.concatenate([q, qp]), np.concatenate([d, plume.increments_at(qp)])
    t_data = time.perf_counter() - t0
    cl = Closure(r, max_batch=max_batch, seed=seed)
    t0 = time.perf_counter()
    losses = cl.fit(q, d, epochs=epochs, batch=512, lr=2e-3, lr_final=1e-5, seed=7,
                    log=(lambda ep, m: log(f"closure epoch {ep}: mse {m:.3e}") if ep % 20 == 0 else None)
                    if log else None)
    return cl, {"samples": int(q.shape[0]), "losses": losses, "data_s": t_data, "train_s": time.perf_counter() - t0}


def fit_scales(q: np.ndarray, d: np.ndarray, floor: float = 1e-12):
    """Per-component input and output scales (1/std of q, std of d)."""
    s_in = 1.0 / np.maximum(q.std(axis=0), floor)
    s_out = np.maximum(d.std(axis=0), floor)
    return s_in.astype(np.float32), s_out.astype(np.float32)


def timed(fn, reps: int, stream: Stream) -> float:
    """Mean milliseconds per call of fn(stream) over reps calls (HIP events)."""
    e0, e1 = Event(), Event()
    fn(stream)
    stream.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn(stream)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps
