#!/usr/bin/env python3
"""Benchmark of the fused compressed-subspace step on one MI355X.

A step decodes every (component, column, block) segment of the basis once and
feeds it to both the project (q = U^T v) and the unproject (field = U q) of
the reference's stepPlume pair (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:477,490).

Workload (default): 128x192x128 grid (peeled 126x190x126, padded 128x192x128,
6144 blocks per component), r = 200 synthetic Fourier modes (16 random terms
each, wavenumbers 1..6, seeded) orthonormalised in FP64 and coded at
percent 0.9995, nBits 32, maxIterations 50; v ~ N(0,1) (seed 42), q ~ N(0,1)
(seed 43).  Inputs are resident in HBM before timing.

Prints one JSON line.  Extra lines go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
WORKLOADS = {
    # name: (full grid, r, percent)
    "128x192x128_r200": ((128, 192, 128), 200, 0.9995),
    "96x128x96_r150": ((96, 128, 96), 150, 0.9995),
    "48x64x48_r52": ((48, 64, 48), 52, 0.9995),
}
DEFAULT = "128x192x128_r200"
BATCH = "batch16"  # 16 independent bases of seeded sizes in one persistent launch
CLOSURE = "closure"  # learned advection closure vs the cubature reduced-advection step, 48x64x48 r=52
PLUME = "plume"  # the whole plume timestep on the device (stepPlume), two bases, HIP-graph replay
PIPELINE = "pipeline"  # full-rank plume -> SVD bases -> reduced operators -> cubature -> compressed step, accuracy
PLUME_CLOSURE = "plume-closure"  # the learned closure inside the real reduced plume, against its cubature
OBSTACLE = "obstacle"  # the moving-obstacle chain: full-rank box run -> bases with U.iop -> reduced obstacle step
ACCURACY = "accuracy"  # compressed vs uncompressed reduced plume, per-step q error, 48x64x48 r=52


def batch_specs(n=16, seed=2024):
    """Seeded grids in [32, 128] x [32, 192] x [32, 128] (multiples of 8),
    r in [20, 200], percent in [0.99, 0.9995]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        grid = (8 * int(rng.integers(4, 17)), 8 * int(rng.integers(4, 25)), 8 * int(rng.integers(4, 17)))
        out.append((grid, int(rng.integers(20, 201)), float(rng.uniform(0.99, 0.9995)), 100 + i))
    return out


def load_or_encode(C, synth, grid, r, percent, seed, name, cache, threads):
    dims = tuple(g - 2 for g in grid)
    paths = [os.path.join(cache, f"U.{name}.component{k}") for k in range(3)] if cache else None
    if paths and all(os.path.exists(p) for p in paths):
        return [open(p, "rb").read() for p in paths]
    U = synth.synthetic_basis(dims, r, seed=seed)
    bufs = C.encode_matrix(U, dims, 32, percent, 50, threads=threads).buffers()
    del U
    if paths:
        os.makedirs(cache, exist_ok=True)
        for p, b in zip(paths, bufs):
            with open(p, "wb") as f:
                f.write(b)
    return bufs


def run_batch(args):
    """All bases of the batch stepped together: one persistent launch plus one
    q reduction per step, autotuned, captured in a HIP graph and replayed."""
    import cfs_plume_amd as C
    from cfs_plume_amd import synth
    from cfs_plume_amd.hip import DeviceArray, Event, Stream, synchronize

    specs = batch_specs()
    t0 = time.perf_counter()
    bases = []
    for i, (grid, r, percent, seed) in enumerate(specs):
        bufs = load_or_encode(C, synth, grid, r, percent, seed,
                              f"batch{i}_{grid[0]}x{grid[1]}x{grid[2]}_r{r}_p{percent:.5f}", args.cache, args.threads)
        bases.append(C.CompressedBasis.from_buffers(bufs))
    log(f"[bench] {len(bases)} bases ready: {time.perf_counter() - t0:.1f}s")
    arrs = []
    for i, b in enumerate(bases):
        v = synth.gaussian_field(b.dims, 42 + i)
        q = synth.gaussian_q(b.r, 43 + i)
        arrs.append((DeviceArray.from_numpy(v.astype(np.float32)), DeviceArray.zeros((b.r,), np.float64),
                     DeviceArray.from_numpy(q), DeviceArray.zeros((len(v),), np.float32)))
    bt = C.BasisBatch(bases)
    bt.bind([a[0] for a in arrs], [a[1] for a in arrs], [a[2] for a in arrs], [a[3] for a in arrs])
    stream = Stream()
    cfg = bt.tune(reps=3, stream=stream)
    log(f"[bench] autotuned launch: {cfg}")
    bt.capture(stream)
    for _ in range(args.warmup):
        bt.replay(stream)
    stream.synchronize()
    for b in bases:
        b.raise_on_stream_errors()
    ev = [Event() for _ in range(args.steps + 1)]
    synchronize()
    ev[0].record(stream)
    for i in range(args.steps):
        bt.replay(stream)
        ev[i + 1].record(stream)
    stream.synchronize()
    per = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(args.steps)])
    ms_step = ev[0].elapsed_ms(ev[-1]) / args.steps
    for b in bases:
        b.raise_on_stream_errors()
    step_bytes = sum(b.step_bytes() for b in bases)
    achieved = step_bytes / (float(per.mean()) * 1e-3) / 1e9
    log(f"[bench] batch of {len(bases)}: {ms_step:.4f} ms/step, B_step={step_bytes / 1e9:.3f} GB -> {achieved:.0f} GB/s")
    # the same work as one fused-step launch per basis, back to back on the stream
    seq_reps = max(1, min(args.steps, 20))
    for b, (v_, qo_, q_, f_) in zip(bases, arrs):
        C.step(b, b, v_, qo_, q_, f_, stream)
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(seq_reps):
        for b, (v_, qo_, q_, f_) in zip(bases, arrs):
            C.step(b, b, v_, qo_, q_, f_, stream)
    e1.record(stream)
    stream.synchronize()
    for b in bases:
        b.raise_on_stream_errors()
    seq_ms = e0.elapsed_ms(e1) / seq_reps
    log(f"[bench] the same {len(bases)} bases one launch each: {seq_ms:.4f} ms/step "
        f"(batched launch {seq_ms / ms_step:.2f}x faster)")
    result = {
        "metric": f"batched reduced steps/s ({len(bases)} bases, decode+project+unproject); % HBM peak",
        "value": round(1000.0 / ms_step, 2), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {
            "workload": f"{len(bases)} independent bases in one persistent launch per step, HIP-graph replay, "
                        f"autotuned launch shape",
            "bases": [{"grid": list(g), "r": r, "percent": round(p, 5)} for g, r, p, _ in specs],
            "bytes_per_step": step_bytes, "launch": cfg, "parallelism": "single GPU",
            "sequential_ms_per_step": round(seq_ms, 5),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_from_profiles(BATCH, "batch_step_kernel")},
    }
    print(json.dumps(result), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps, warmup, stream):
    """`warmup` untimed calls, then `steps` calls with a HIP event recorded on
    `stream` after each: (mean ms per call over the whole region, per-call ms)."""
    from cfs_plume_amd.hip import Event, synchronize
    for _ in range(warmup):
        fn()
    stream.synchronize()
    ev = [Event() for _ in range(steps + 1)]
    synchronize()
    ev[0].record(stream)
    for i in range(steps):
        fn()
        ev[i + 1].record(stream)
    stream.synchronize()
    synchronize()
    per = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(steps)])
    return ev[0].elapsed_ms(ev[-1]) / steps, per


def run_closure(args):
    """Learned closure against the plain cubature step on the 48x64x48,
    r = 52 basis.  A cubature (key cells and weights, generateImportanceSampledCubature)
    is fitted to the exact reduced advection U^T A(U q) at states of exact
    runs; the reduced plume stepped with it is the plain solver, whose
    trajectories train the MLP on the device.  Held-out runs are rolled out
    with the closure and with the cubature step (the judge), the cubature's
    own error is measured against the exact step, and one step of each is
    timed (n = 1, HIP events)."""
    import cfs_plume_amd as C
    from cfs_plume_amd import closure as CL
    from cfs_plume_amd import synth
    from cfs_plume_amd.hip import DeviceArray, Stream

    grid, r, percent = (48, 64, 48), 52, 0.9995
    n_train, n_held, steps = args.closure_runs, 20, 100
    bufs = load_or_encode(C, synth, grid, r, percent, 7, "48x64x48_r52", args.cache, args.threads)
    U = C.CompressedBasis.from_buffers(bufs)
    plume = C.ReducedPlume(U, dt0=1.0, alpha=0.99, force=25.0)
    runs = C.ReducedPlume.seeded_runs(n_train + n_held, seed=2026)
    b_train = plume.forcing(runs[:n_train])
    b_held = plume.forcing(runs[n_train:])

    # the cubature: exact reduced advection at states of exact training runs,
    # half on the trajectories and half perturbed around them
    t0 = time.perf_counter()
    if args.cubature_file:
        # a cubature written earlier (e.g. the committed fixture of this basis)
        from cfs_plume_amd import cubature as CB
        plume.cubature = CB.Cubature.from_file(args.cubature_file, U)
        cells, weights = plume.cubature.cells_and_weights()
        res = CB.CubatureResult(cells, weights, float("nan"), 0)
        states = np.zeros((0, r))
    else:
        states = CL.cubature_training_states(plume, b_train, args.cubature_runs, steps, args.cubature_samples,
                                             args.closure_noise)
        res = plume.make_cubature(states, candidates_per_pass=args.cubature_candidates, passes=6,
                                  tolerance=args.cubature_tolerance, log=lambda m: log(f"[bench] {m}"))
    t_cub = time.perf_counter() - t0
    exact = plume.rollout_exact(b_held, n_held, steps)
    cub = plume.rollout(b_held, n_held, steps)

    def drift(a, ref):
        e = np.linalg.norm(a[:, 1:] - ref[:, 1:], axis=2) / np.linalg.norm(ref[:, 1:], axis=2)
        return float(e[:, -1].mean()), float(e.max())

    d_cub, d_cub_max = drift(cub, exact)
    log(f"[bench] cubature: {len(res.cells)} key cells after {res.passes} passes ({t_cub:.1f}s), training error "
        f"{res.relative_error:.3e}; held-out drift against the exact step at step {steps}: {d_cub:.3e}")

    cl, info = CL.train_closure(plume, b_train, n_train, steps, args.closure_epochs, args.closure_noise,
                                log=lambda m: log(f"[bench] {m}"))
    losses = info["losses"]
    log(f"[bench] closure: {info['samples']} samples of the cubature solver in {info['data_s']:.1f}s, "
        f"{args.closure_epochs} epochs in {info['train_s']:.1f}s, final mse {losses[-1]:.3e} (first {losses[0]:.3e})")
    learned = plume.rollout_closure(cl, b_held, n_held, steps)
    bh = b_held.numpy()
    plain = np.zeros_like(cub)
    for t in range(steps):
        plain[:, t + 1] = plume.alpha * plain[:, t] + bh
    d_cl, d_cl_max = drift(learned, cub)
    d_cl_exact, _ = drift(learned, exact)
    d_pl, _ = drift(plain, cub)
    checked, falls = plume.rollout_closure_checked(cl, b_held, 0, steps, check_every=10, tol=0.01)
    d_chk = float(np.linalg.norm(checked[-1] - cub[0, -1]) / np.linalg.norm(cub[0, -1]))
    log(f"[bench] held-out drift against the cubature step at step {steps}: closure {d_cl:.3e} (max over steps "
        f"{d_cl_max:.3e}), without advection {d_pl:.3e}; checked every 10 steps: {d_chk:.3e} with {falls} fallbacks")

    # one step of each, n = 1
    stream = Stream()
    qa = DeviceArray.from_numpy(cub[0, steps // 2])
    qb = DeviceArray.zeros((r,), np.float64)
    reps = args.steps
    ms_exact = CL.timed(lambda s: C.reduced_advect(U, qa, qb, plume.fa, plume.fb, plume.dt0, plume.alpha,
                                                   (b_held, 0), s), reps, stream)
    ms_cub = CL.timed(lambda s: plume.advect(qa, qb, plume.alpha, (b_held, 0), s), reps, stream)
    ms_cl = CL.timed(lambda s: cl.step(qa, qb, 1, plume.alpha, (b_held, 0), s), reps, stream)
    log(f"[bench] step time: closure {ms_cl * 1e3:.1f} us, cubature {ms_cub * 1e3:.1f} us "
        f"({ms_cub / ms_cl:.1f}x), exact reduced advection {ms_exact * 1e3:.1f} us ({ms_exact / ms_cl:.1f}x)")
    result = {
        "metric": "closure reduced-advection steps/s (r=52 @48x64x48); speedup over the cubature step",
        "value": round(1000.0 / ms_cl, 1), "unit": "steps/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": round(ms_cl, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 weights, fp32 accumulation", "data": "synthetic",
        "config": {
            "workload": f"learned advection closure (MLP {r}-256-256-256-{r}, GELU) on the grid 48x64x48, r={r}, "
                        f"against the cubature step; {n_train} training runs and {n_held} held-out runs of "
                        f"{steps} steps",
            "grid": list(grid), "r": r, "parallelism": "single GPU",
        },
        "cubature": {
            "key_cells": int(len(res.cells)), "passes": res.passes,
            "training_rel_error": None if args.cubature_file else res.relative_error,
            "file": os.path.basename(args.cubature_file) if args.cubature_file else None,
            "samples": int(len(states)), "candidates_per_pass": args.cubature_candidates,
            "tolerance": args.cubature_tolerance,
            "fit_s": round(t_cub, 2), "drift_vs_exact_at_100": round(d_cub, 6),
            "drift_vs_exact_max": round(d_cub_max, 6), "step_ms": round(ms_cub, 5),
        },
        "closure": {
            "cubature_step_ms": round(ms_cub, 5), "speedup_over_cubature": round(ms_cub / ms_cl, 2),
            "exact_step_ms": round(ms_exact, 5), "speedup_over_exact": round(ms_exact / ms_cl, 2),
            "drift_rel_l2_at_100": round(d_cl, 6), "drift_rel_l2_max": round(d_cl_max, 6),
            "drift_vs_exact_at_100": round(d_cl_exact, 6),
            "drift_without_advection": round(d_pl, 6),
            "checked_every_10": {"drift_at_100": round(d_chk, 6), "fallbacks": falls},
            "train": {"samples": info["samples"], "epochs": args.closure_epochs, "batch": 512,
                      "perturbed_states": args.closure_noise, "linear_skip": "least squares, fp32",
                      "mse_first": losses[0], "mse_last": losses[-1], "data_s": round(info["data_s"], 2),
                      "train_s": round(info["train_s"], 2)},
        },
    }
    print(json.dumps(result), flush=True)


def run_plume(args):
    """The timestep an integrator runs (addSmokeColumn + stepPlume): forces,
    project with U_preadvect, density/heat advection, the subspace advection,
    reduced diffusion and pressure projection, unproject with U_final -- two
    passes over two different compressed bases, captured in one HIP graph.
    Default grid 128x192x128, r = 200 (both bases synthetic, seeds 7 and 8)
    with the exact reduced advection U^T A(U q) (two more passes); at
    48x64x48, r = 52 also with the committed cubature fixture."""
    import cfs_plume_amd as C
    from cfs_plume_amd import cubature as CB
    from cfs_plume_amd import plume as PL
    from cfs_plume_amd import synth
    from cfs_plume_amd.hip import DeviceArray, Stream

    name = args.plume_grid
    grid, r, percent = WORKLOADS[name]
    pre = C.CompressedBasis.from_buffers(load_or_encode(C, synth, grid, r, percent, 7, name, args.cache,
                                                        args.threads))
    fin = C.CompressedBasis.from_buffers(load_or_encode(C, synth, grid, r, percent, 8, name + "_fin", args.cache,
                                                        args.threads))
    rng = np.random.default_rng(5)
    D = 0.97 * np.eye(r) + 0.01 * rng.standard_normal((r, r)) / np.sqrt(r)
    P = np.eye(r) + 0.01 * rng.standard_normal((r, r)) / np.sqrt(r)
    I = 0.02 * rng.standard_normal((r, r)) / np.sqrt(r)
    solver = PL.PlumeSolver(pre, fin, D, P, I)
    modes = [("exact", PL.ADVECT_EXACT, None)]
    fixture = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "cubature_48x64x48_r52")
    if name == "48x64x48_r52" and os.path.exists(fixture):
        modes.append(("cubature", PL.ADVECT_CUBATURE, CB.Cubature.from_file(fixture, pre)))
    stream = Stream()
    out = {}
    for label, mode, cub in modes:
        solver.set_advection(mode, cubature=cub)
        vel, dens, heat, q = solver.fields()
        for _ in range(3):
            solver.frame(vel, dens, heat, q, stream)
        graph = solver.capture(vel, dens, heat, q, stream)
        ms, per = time_launches(graph.launch, args.steps, args.warmup, stream)
        pre.raise_on_stream_errors()
        qn = q.numpy()
        out[label] = {"ms_per_frame": round(ms, 5), "median_ms": round(float(np.median(per)), 5),
                      "frames_per_s": round(1000.0 / ms, 1), "q_norm": float(np.linalg.norm(qn))}
        log(f"[bench] plume frame ({label} advection), {name}: {ms:.4f} ms (median {np.median(per):.4f}), "
            f"|q| {np.linalg.norm(qn):.3e}")
    # the step's two compressed passes alone
    dv = DeviceArray.zeros((3 * solver.full[0] * solver.full[1] * solver.full[2],), np.float32)
    qa = DeviceArray.from_numpy(rng.standard_normal(r))
    qb = DeviceArray.zeros((r,), np.float64)
    ms_p, _ = time_launches(lambda: pre.project(dv, qb, stream), args.steps, args.warmup, stream)
    ms_u, _ = time_launches(lambda: fin.unproject(qa, dv, stream), args.steps, args.warmup, stream)
    first = modes[0][0]
    result = {
        "metric": f"plume timesteps/s (addSmokeColumn + stepPlume on the device, {first} advection), "
                  f"r={r} @{grid[0]}x{grid[1]}x{grid[2]}",
        "value": out[first]["frames_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out[first]["ms_per_frame"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 fields, fp64 q", "data": "synthetic",
        "config": {
            "workload": f"stepPlume with U_preadvect != U_final (synthetic, seeds 7 and 8), grid "
                        f"{grid[0]}x{grid[1]}x{grid[2]}, r={r}, percent={percent}; one HIP graph per frame",
            "grid": list(grid), "r": r, "parallelism": "single GPU",
        },
        "frames": out,
        "passes_ms": {"project_pre": round(ms_p, 5), "unproject_fin": round(ms_u, 5)},
    }
    solver.close()
    print(json.dumps(result), flush=True)


def run_obstacle(args):
    """The moving-obstacle chain on one MI355X: the full-rank solver stepped
    with the spinning, sliding box (stepWithMovingObstacle, the box of
    buildProducts.cpp: centre 0.5, lengths 0.4 x 0.05 x 0.15, periods 4 and
    10), SVD bases including U.iop, the IOP products, the staged cubature and
    the compressed bases, then the reduced obstacle step (stepMovingObstacle,
    the box's R q + n evaluated from its cells) with the compressed and the
    uncompressed bases against the full-rank run; one reduced obstacle frame
    timed (compressed bases, cubature advection)."""
    from cfs_plume_amd import accuracy as A
    from cfs_plume_amd import obstacle as OBX
    from cfs_plume_amd import pipeline as PP
    from cfs_plume_amd.cubature import Cubature
    from cfs_plume_amd.hip import Stream
    from cfs_plume_amd.plume import ADVECT_CUBATURE, PlumeSolver, default_params

    grid = tuple(int(g) for g in args.pipeline_grid.split("x"))
    steps = args.pipeline_steps
    t0 = time.perf_counter()
    data = PP.build_plume_data(grid, steps, cubature_snapshots=args.pipeline_cubature_snapshots,
                               candidates=args.pipeline_candidates, tolerance=args.pipeline_tolerance,
                               threads=args.threads, obstacle=OBX.Box(), log=lambda m: log(f"[bench] {m}"))
    t_build = time.perf_counter() - t0
    ranks = {k: b.r for k, b in data.bases.items()}
    log(f"[bench] obstacle chain built in {t_build:.1f}s: ranks {ranks}, cubature {len(data.cubature.cells)} cells, "
        f"timings {data.timings}")
    t0 = time.perf_counter()
    acc = A.compare_plume(data, steps)
    t_runs = time.perf_counter() - t0
    e = acc["q_rel_l2"]
    last = {name: round(float(acc[name]["velocity_vs_full_rank"][-1]), 8) for name in
            ("compressed", "uncompressed", "uncompressed_exact") if name in acc}
    log(f"[bench] obstacle: compressed vs uncompressed q max {e.max():.3e}; velocity vs full rank at the last step "
        f"{last} ({t_runs:.1f}s)")
    # one reduced obstacle frame (compressed bases, cubature), the box held where it starts
    ops = data.operators
    dims = data.peeled
    comp = PP.compressed_bases(data)
    prm = default_params(data.full)
    prm.vorticity_eps = 0.0
    zeros = np.zeros((data.bases["final"].r, ops["damping.peeled"].shape[0]))
    solver = PlumeSolver(comp["preadvect"], comp["final"], ops["damping.peeled"], zeros, zeros, prm)
    cub = Cubature.from_rows(data.cubature.cells, data.cubature.weights, dims, data.bases["prediffuse"].numpy())
    solver.set_advection(ADVECT_CUBATURE, cubature=cub)
    solver.set_obstacle(data.bases["iop"].U, data.bases["preproject"].U, ops["U'U.iop.subspace"],
                        ops["projected.ptof"], ops["inverseProduct"])
    box = OBX.Box()
    stream = Stream()
    vel, dens, heat, q = solver.fields()
    for _ in range(10):
        solver.frame_obstacle(box, vel, dens, heat, q, stream)
    ms, per = time_launches(lambda: solver.step_obstacle(box, vel, dens, heat, q, stream), args.steps, args.warmup,
                            stream)
    full_ms = 1e3 * data.timings["full_rank_s"] / steps
    result = {
        "metric": f"reduced moving-obstacle frame time (stam grid {grid[0]}x{grid[1]}x{grid[2]}, compressed bases, "
                  f"cubature advection)",
        "value": round(ms, 5), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 5), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64 full-rank and bases, fp32 reduced fields, fp64 q",
        "data": "simulated (full-rank moving-obstacle run on the device)",
        "config": {"workload": f"stepWithMovingObstacle x {steps} -> svdOutOfCoreMultiple (iop) -> "
                               f"buildProducts (iop) -> cubature -> compressMatrices -> stepMovingObstacle",
                   "grid": list(grid), "ranks": ranks, "parallelism": "single GPU",
                   "box": {"center": [0.5, 0.5, 0.5], "lengths": [0.4, 0.05, 0.15], "period": 4.0,
                           "translation_period": 10.0}},
        "accuracy": {"q_rel_l2_compressed_vs_uncompressed_max": round(float(e.max()), 6),
                     "q_rel_l2_at": {str(s): round(float(e[s - 1]), 6) for s in (1, 10, 50, 100) if s <= steps},
                     "velocity_vs_full_rank_last": last},
        "full_rank_ms_per_step": round(full_ms, 3), "speedup_over_full_rank": round(full_ms / ms, 1),
        "median_ms": round(float(np.median(per)), 5),
        "timings_s": {k: round(v, 2) for k, v in data.timings.items()} | {"reduced_runs": round(t_runs, 2)},
    }
    solver.close()
    cub.close()
    print(json.dumps(result), flush=True)


def run_pipeline(args):
    """The reference's whole offline chain on one MI355X (cfg stam.64: a
    48x64x48 Stam plume, 100 full-rank steps, SVD bases with the 1e-9 keep
    rule, buildProducts, the staged cubature from 20 snapshot pairs, the
    bases compressed at percent 0.9995, nBits 32; the cubature trained on all
    100 snapshot pairs), then the reduced plume
    stepped 100 frames with the compressed and with the uncompressed bases
    (compareSubspace per step, the measurement behind the README's
    0.25 / 0.91 / 1.06 % at steps 1 / 50 / 100), and the fused step timed on
    the plume's own compressed U_final."""
    import cfs_plume_amd as C
    from cfs_plume_amd import accuracy as A
    from cfs_plume_amd import pipeline as PP
    from cfs_plume_amd.hip import DeviceArray, Stream

    grid = tuple(int(g) for g in args.pipeline_grid.split("x"))
    steps = args.pipeline_steps
    t0 = time.perf_counter()
    data = PP.build_plume_data(grid, steps, cubature_snapshots=args.pipeline_cubature_snapshots,
                               candidates=args.pipeline_candidates, tolerance=args.pipeline_tolerance,
                               threads=args.threads, log=lambda m: log(f"[bench] {m}"))
    t_build = time.perf_counter() - t0
    ranks = {k: b.r for k, b in data.bases.items()}
    log(f"[bench] pipeline built in {t_build:.1f}s: ranks {ranks}, cubature {len(data.cubature.cells)} cells "
        f"(error {data.cubature.relative_error:.3e}), timings {data.timings}")
    if args.pipeline_out:
        PP.write_reduced(data, args.pipeline_out)
    t0 = time.perf_counter()
    acc = A.compare_plume(data, steps, out_prefix=os.path.join(args.pipeline_out, "run") if args.pipeline_out else None)
    t_runs = time.perf_counter() - t0
    e = acc["q_rel_l2"]
    pick = {str(s): round(float(e[s - 1]), 6) for s in (1, 10, 50, 100) if s <= steps}
    vc, vu = acc["compressed"]["velocity_vs_full_rank"], acc["uncompressed"]["velocity_vs_full_rank"]
    for name in ("compressed", "uncompressed", "uncompressed_exact"):
        if name in acc:
            v = acc[name]["velocity_vs_full_rank"]
            log(f"[bench] {name} velocity vs full rank: " +
                " ".join(f"{s}:{v[s - 1]:.3e}" for s in (1, 2, 5, 10, 20, 50, 100) if s <= len(v)))
    log(f"[bench] compressed vs uncompressed q: {pick} (max {e.max():.3e}); velocity vs full rank at the last "
        f"step: uncompressed {vu[-1]:.3e}, compressed {vc[-1]:.3e} ({t_runs:.1f}s)")
    # the fused step on the plume's own compressed U_final
    U = C.CompressedBasis.from_buffers(data.compressed["final"])
    r = U.r
    stream = Stream()
    rng = np.random.default_rng(42)
    n3 = 3 * (grid[0] * grid[1] * grid[2])
    dv = DeviceArray.from_numpy(rng.standard_normal(n3).astype(np.float32))
    dq = DeviceArray.from_numpy(rng.standard_normal(r))
    qo = DeviceArray.zeros((r,), np.float64)
    fo = DeviceArray.zeros((n3,), np.float32)
    ms, per = time_launches(lambda: C.step(U, U, dv, qo, dq, fo, stream), args.steps, args.warmup, stream)
    U.raise_on_stream_errors()
    step_bytes = U.step_bytes()
    result = {
        "metric": f"compressed vs uncompressed reduced q, relative L2 at step {steps} (stam plume "
                  f"{grid[0]}x{grid[1]}x{grid[2]}, percent 0.9995)",
        "value": round(float(e[-1]), 6), "unit": "relative L2", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp64 full-rank and bases, fp32 reduced fields, fp64 q", "data": "simulated (full-rank plume on the device)",
        "config": {
            "workload": f"fluid3D -> svdOutOfCoreMultiple -> buildProducts -> cubatureGeneratorStamStaged -> "
                        f"compressMatrices -> reducedCompressed3D vs reduced3D, {steps} frames",
            "grid": list(grid), "ranks": ranks, "parallelism": "single GPU",
        },
        "accuracy": {"q_rel_l2": pick, "q_rel_l2_max": round(float(e.max()), 6),
                     "reference_readme_rel_l2": {"1": 0.0025, "50": 0.0091, "100": 0.0106},
                     "velocity_vs_full_rank_last": {"uncompressed": round(float(vu[-1]), 6),
                                                    "compressed": round(float(vc[-1]), 6)}},
        "cubature": {"key_cells": int(len(data.cubature.cells)), "training_rel_error": data.cubature.relative_error,
                     "passes": data.cubature.passes},
        "fused_step_on_plume_basis": {"ms": round(ms, 5), "median_ms": round(float(np.median(per)), 5),
                                      "r": r, "bytes": step_bytes,
                                      "GB_per_s": round(step_bytes / (ms * 1e-3) / 1e9, 1)},
        "timings_s": {k: round(v, 2) for k, v in data.timings.items()} | {"reduced_runs": round(t_runs, 2)},
    }
    print(json.dumps(result), flush=True)


def run_plume_closure(args):
    """C5 on the plume's own bases: the pipeline's 48x64x48 plume (compressed
    U_preadvect r = 52, U_final r = 46, the staged cubature into U_prediffuse
    coordinates r = 53); 200 seeded reduced-plume runs of 100 frames (source
    offsets uniform in +-8 cells, buoyancy 0.5-2x) with the cubature train
    the MLP (r = 53, 3 x 256, GELU) on the device; 20 held-out runs with the
    closure inside stepPlume are compared with their cubature runs, and the
    advection stage is timed both ways."""
    from cfs_plume_amd import pipeline as PP
    from cfs_plume_amd import plume_closure as PC

    grid = tuple(int(g) for g in args.pipeline_grid.split("x"))
    data = PP.build_plume_data(grid, args.pipeline_steps, cubature_snapshots=args.pipeline_cubature_snapshots,
                               candidates=args.pipeline_candidates, tolerance=args.pipeline_tolerance,
                               threads=args.threads, log=lambda m: log(f"[bench] {m}"))
    ex = PC.PlumeClosureExperiment(data, frames=100)
    runs = PC.seeded_plume_runs(220, seed=2026)
    cl, info = ex.train(runs[:200], epochs=args.plume_closure_epochs, noise=args.plume_closure_noise,
                        log=lambda m: log(f"[bench] {m}"))
    log(f"[bench] plume closure: {info['samples']} samples in {info['data_s']:.1f}s, trained in "
        f"{info['train_s']:.1f}s, mse {info['losses'][0]:.3e} -> {info['losses'][-1]:.3e}")
    ev = ex.evaluate(cl, runs[200:])
    d = ev["drift"]
    times = ex.step_times(cl, args.steps)
    c, o = ev["checked_drift"], ev["one_step"]
    log(f"[bench] checked every 10 frames: step 100 mean {c[:, -1].mean():.3e} (max {c[:, -1].max():.3e}), "
        f"{ev['fallback_frames']} fallback frames of {d.size}; one-step error mean {o.mean():.3e} "
        f"(median {np.median(o):.3e}, max {o.max():.3e})")
    log(f"[bench] held-out drift against the cubature runs: step 100 mean {d[:, -1].mean():.3e} (max "
        f"{d[:, -1].max():.3e}), worst frame mean {d.mean(axis=0).max():.3e}; advection stage: closure "
        f"{times['closure_ms'] * 1e3:.1f} us, cubature {times['cubature_ms'] * 1e3:.1f} us "
        f"({times['cubature_ms'] / times['closure_ms']:.1f}x)")
    result = {
        "metric": "learned closure inside the reduced plume: advection-stage speedup over the cubature "
                  "(r_pre=52 -> r_adv=53 @48x64x48 plume basis)",
        "value": round(times["cubature_ms"] / times["closure_ms"], 2), "unit": "x", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(times["closure_ms"], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 weights, fp32 accumulation",
        "data": "simulated (full-rank plume -> reduced plume runs on the device)",
        "config": {"workload": "C5 on the plume's own bases: 200 training runs x 100 frames, 20 held out",
                   "grid": list(grid), "parallelism": "single GPU"},
        "closure": {"drift_at_100_mean": round(float(d[:, -1].mean()), 6),
                    "drift_at_100_max": round(float(d[:, -1].max()), 6),
                    "drift_worst_frame_mean": round(float(d.mean(axis=0).max()), 6),
                    "checked_every_10": {"drift_at_100_mean": round(float(c[:, -1].mean()), 6),
                                         "fallback_frames": int(ev["fallback_frames"]), "frames": int(d.size)},
                    "one_step_error": {"mean": round(float(o.mean()), 6), "median": round(float(np.median(o)), 6),
                                       "max": round(float(o.max()), 6)},
                    "closure_ms": round(times["closure_ms"], 5), "cubature_ms": round(times["cubature_ms"], 5),
                    "samples": info["samples"], "epochs": args.plume_closure_epochs,
                    "perturbed_states": args.plume_closure_noise,
                    "mse_first": info["losses"][0], "mse_last": info["losses"][-1],
                    "data_s": round(info["data_s"], 2), "train_s": round(info["train_s"], 2)},
    }
    ex.close()
    print(json.dumps(result), flush=True)


def run_accuracy(args):
    """Compression error of the reduced plume (the measurement of the
    reference's reducedCompressed3DCompareVelocities): the 48x64x48, r = 52
    basis coded at percent 0.9995, seeded runs stepped 100 times with the
    compressed basis on the device and with the dense basis it was encoded
    from; per-step relative L2 error of q."""
    import cfs_plume_amd as C
    from cfs_plume_amd import accuracy as A
    from cfs_plume_amd import synth

    grid, r, percent, steps = (48, 64, 48), 52, 0.9995, 100
    dims = tuple(g - 2 for g in grid)
    U_dense = synth.synthetic_basis(dims, r, seed=7)
    bufs = load_or_encode(C, synth, grid, r, percent, 7, "48x64x48_r52", args.cache, args.threads)
    U = C.CompressedBasis.from_buffers(bufs)
    plume = C.ReducedPlume(U, dt0=1.0, alpha=0.99, force=25.0)
    runs = C.ReducedPlume.seeded_runs(args.accuracy_runs, seed=2026)
    t0 = time.perf_counter()
    errs = A.compare_runs(plume, U_dense, runs, steps, out_prefix=args.accuracy_out or None)
    mean = errs.mean(axis=0)
    log(f"[bench] compressed vs uncompressed q over {len(runs)} runs ({time.perf_counter() - t0:.1f} s): "
        f"step 1 {mean[0]:.3%}, step 50 {mean[49]:.3%}, step 100 {mean[99]:.3%}, max {errs.max():.3%}")
    result = {
        "metric": "compressed vs uncompressed reduced q, relative L2 at step 100 (r=52 @48x64x48, percent 0.9995)",
        "value": round(float(mean[steps - 1]), 6), "unit": "relative L2", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 fields, fp64 q", "data": "synthetic",
        "config": {
            "workload": f"reduced plume stand-in q' = 0.99 U^T A(U q) + U^T f, {len(runs)} seeded runs of {steps} "
                        f"steps, compressed (device) against dense FP64 basis",
            "grid": list(grid), "r": r, "percent": percent, "parallelism": "single GPU",
        },
        "accuracy": {
            "mean_rel_l2": {"1": round(float(mean[0]), 6), "10": round(float(mean[9]), 6),
                            "50": round(float(mean[49]), 6), "100": round(float(mean[99]), 6)},
            "max_rel_l2": round(float(errs.max()), 6),
            "reference_readme_rel_l2": {"1": 0.0025, "50": 0.0091, "100": 0.0106},
        },
    }
    print(json.dumps(result), flush=True)


def traffic_from_profiles(workload: str, kernel: str = "step_kernel"):
    """HBM bytes per launch of `kernel` from the committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def mfma_mops_from_profiles(workload: str):
    """SQ_INSTS_VALU_MFMA_MOPS_BF16 per step launch (units of 512 FLOP) from the
    committed PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", f"step_kernel_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"SQ_INSTS_VALU_MFMA_MOPS_BF16\s*\|\s*([0-9,]+)", open(path).read())
    return float(m.group(1).replace(",", "")) if m else None


def host_cpu():
    """The host CPU's model name, its logical CPUs and those this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(bufs, v, q, r, sample_cols):
    """The FP64 oracle (single thread) on columns [0, sample_cols); a partial
    sample is extrapolated linearly in the column count to the full step."""
    from oracle import oracle as O
    ob = O.Basis(buffers=bufs)
    t0 = time.perf_counter()
    ob.project(v, 0, 0)
    ob.unproject(np.zeros(0), v, 0, 0)
    t_fixed = time.perf_counter() - t0
    t0 = time.perf_counter()
    q_s = ob.project(v, 0, sample_cols)
    f_s = ob.unproject(q[:sample_cols], np.zeros_like(v), 0, sample_cols)
    t_sample = time.perf_counter() - t0
    t_full = t_sample if sample_cols >= r else t_fixed + (t_sample - t_fixed) * r / sample_cols
    ob.close()
    return 1.0 / t_full, t_full, t_sample, q_s, f_s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS) + [BATCH, CLOSURE, ACCURACY, PLUME, PIPELINE, PLUME_CLOSURE, OBSTACLE])
    ap.add_argument("--pipeline-grid", default="48x64x48", help="full grid of the pipeline workload")
    ap.add_argument("--pipeline-steps", type=int, default=100, help="full-rank steps and reduced frames")
    ap.add_argument("--pipeline-cubature-snapshots", type=int, default=100,
                    help="snapshot pairs the cubature is trained on (cfg 'reduced snapshots'; the default 20 "
                         "of the reference's parser trains on the early plume only, and the reduced run diverges)")
    ap.add_argument("--pipeline-candidates", type=int, default=1000, help="cubature candidates per pass")
    ap.add_argument("--pipeline-tolerance", type=float, default=0.01, help="cubature error that ends the passes")
    ap.add_argument("--pipeline-out", default="", help="write the reduced files and q trajectories here")
    ap.add_argument("--plume-closure-epochs", type=int, default=3000)
    ap.add_argument("--plume-closure-noise", type=float, default=0.02,
                    help="relative size of the perturbed states of the plume-closure workload")
    ap.add_argument("--plume-grid", default=DEFAULT, choices=sorted(WORKLOADS),
                    help="grid and r of the plume workload")
    ap.add_argument("--accuracy-runs", type=int, default=4, help="seeded runs of the accuracy workload")
    ap.add_argument("--accuracy-out", default="",
                    help="write each accuracy run's trajectories as <prefix>.<i>[.compressed].qDot.matrix")
    ap.add_argument("--closure-runs", type=int, default=200, help="training runs of the closure workload")
    ap.add_argument("--closure-epochs", type=int, default=600)
    ap.add_argument("--cubature-samples", type=int, default=160,
                    help="exact reduced-advection states the closure workload's cubature is fitted to")
    ap.add_argument("--cubature-runs", type=int, default=120, help="exact runs the cubature's states come from")
    ap.add_argument("--cubature-candidates", type=int, default=2000, help="candidate key cells per generator pass")
    ap.add_argument("--cubature-tolerance", type=float, default=1e-3, help="relative fit error that ends the passes")
    ap.add_argument("--cubature-file", default="",
                    help="read the closure workload's cubature from this file instead of fitting one")
    ap.add_argument("--closure-noise", type=float, default=0.2,
                    help="relative size of the perturbed states added to the closure's training data")
    ap.add_argument("--cpu-sample-cols", type=int, default=200,
                    help="columns of the oracle CPU step to time (the full step when >= r)")
    ap.add_argument("--settle-ms", type=float, default=1500.0,
                    help="untimed back-to-back steps before timing (default workload), wall-clock ms")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-spatial", action="store_true", help="skip the spatial decode of all columns")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--cache", default="", help="directory to keep/reuse the encoded component files")
    ap.add_argument("--mode", default="step", choices=["step", "project", "unproject"],
                    help="time the fused step (default) or one half of it")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU split is a design "
              "note (blocks sharded across GPUs, one r-length all-reduce per step) and is not built.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == BATCH:
        run_batch(args)
        return
    if args.workload == CLOSURE:
        run_closure(args)
        return
    if args.workload == ACCURACY:
        run_accuracy(args)
        return
    if args.workload == PLUME:
        run_plume(args)
        return
    if args.workload == OBSTACLE:
        run_obstacle(args)
        return
    if args.workload == PIPELINE:
        run_pipeline(args)
        return
    if args.workload == PLUME_CLOSURE:
        run_plume_closure(args)
        return

    import cfs_plume_amd as C
    from cfs_plume_amd import synth
    from cfs_plume_amd.hip import DeviceArray, Event, Stream, synchronize

    grid, r, percent = WORKLOADS[args.workload]
    dims = tuple(g - 2 for g in grid)
    cache = [os.path.join(args.cache, f"U.{args.workload}.component{k}") for k in range(3)] if args.cache else None
    if cache and all(os.path.exists(p) for p in cache):
        bufs = [open(p, "rb").read() for p in cache]
        log(f"[bench] reusing encoded basis from {args.cache}")
    else:
        t0 = time.perf_counter()
        U = synth.synthetic_basis(dims, r, seed=7)
        log(f"[bench] synthetic basis {dims} r={r}: {time.perf_counter() - t0:.1f}s")
        t0 = time.perf_counter()
        enc = C.encode_matrix(U, dims, 32, percent, 50, threads=args.threads)
        log(f"[bench] encoded: {time.perf_counter() - t0:.1f}s")
        del U
        bufs = enc.buffers()
        del enc
        if cache:
            os.makedirs(args.cache, exist_ok=True)
            for p, b in zip(cache, bufs):
                with open(p, "wb") as f:
                    f.write(b)
    t0 = time.perf_counter()
    basis = C.CompressedBasis.from_buffers(bufs)
    tokens = basis.tokens
    log(f"[bench] opened on device: {time.perf_counter() - t0:.1f}s tokens={tokens} "
        f"mean T/segment={sum(tokens) / (3 * r * basis.nB):.1f} device={basis.device_bytes / 1e9:.2f} GB")

    v = synth.gaussian_field(dims, 42)
    q = synth.gaussian_q(r, 43)
    dv = DeviceArray.from_numpy(v.astype(np.float32))
    dq = DeviceArray.from_numpy(q)
    qo = DeviceArray.zeros((r,), np.float64)
    fo = DeviceArray.zeros((len(v),), np.float32)
    stream = Stream()

    fused = lambda: C.step(basis, basis, dv, qo, dq, fo, stream)  # noqa: E731
    project = lambda: basis.project(dv, qo, stream)  # noqa: E731
    unproject = lambda: basis.unproject(dq, fo, stream)  # noqa: E731
    run_once = {"step": fused, "project": project, "unproject": unproject}[args.mode]

    # Spatial decode of every column (DecodeVectorField for c = 0..r-1): the
    # inverse DCTs run on MFMA (BF16 hi + lo products, FP32 accumulation).
    # Timed before the step; its per-column error is checked further down.
    spatial = None
    cols = None
    if not args.no_spatial:
        cols = DeviceArray((r, 3 * basis.peeled_cells), np.float32)
        basis.decode_columns(0, r, cols, stream)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(3):
            basis.decode_columns(0, r, cols, stream)
        e1.record(stream)
        stream.synchronize()
        basis.raise_on_stream_errors()
        ms_dec = e0.elapsed_ms(e1) / 3
        out_bytes = r * 3 * basis.peeled_cells * 4
        in_bytes = 4 * sum(tokens) + 24 * 3 * r * basis.nB
        spatial = {
            "columns": r, "ms": round(ms_dec, 4), "output_GB": round(out_bytes / 1e9, 3),
            "GB_per_s": round((out_bytes + in_bytes) / (ms_dec * 1e-3) / 1e9, 1),
            "inverse_dct": "MFMA bf16 hi+lo split products, fp32 accumulation",
        }
        log(f"[bench] spatial decode of {r} columns: {ms_dec:.3f} ms ({spatial['GB_per_s']} GB/s read+written)")

    # Settle the GPU before anything is timed: fused steps back to back for
    # --settle-ms of wall time (untimed).  Measured on a cold GPU, launch
    # times climb for the first few dozen launches and settle only after a
    # few hundred (clock and power management), so a short timed run would
    # otherwise measure the ramp rather than the kernel.
    settle = 0
    t_settle = time.perf_counter()
    while (time.perf_counter() - t_settle) * 1e3 < args.settle_ms:
        for _ in range(50):
            run_once()
        stream.synchronize()
        settle += 50
    if settle:
        basis.raise_on_stream_errors()
        log(f"[bench] settled: {settle} untimed launches in {time.perf_counter() - t_settle:.2f} s")

    # The two halves of the step on their own (an integrator's step runs
    # project with U_preadvect and unproject with U_final, two passes).
    halves = {}
    if args.mode == "step":
        for name, fn in (("project", project), ("unproject", unproject)):
            ms_h, per_h = time_launches(fn, args.steps, args.warmup, stream)
            halves[f"{name}_ms"] = round(ms_h, 5)
            log(f"[bench] {name} alone: {ms_h:.4f} ms/launch (median {np.median(per_h):.4f})")
        halves["sum_ms"] = round(halves["project_ms"] + halves["unproject_ms"], 5)
        basis.raise_on_stream_errors()

    # the timed region: W untimed steps, then K steps bracketed by HIP events
    # on the step's stream (one event per launch)
    t_wall = time.perf_counter()
    ms_step, per = time_launches(run_once, args.steps, args.warmup, stream)
    t_wall = time.perf_counter() - t_wall
    basis.raise_on_stream_errors()
    step_bytes = basis.step_bytes()
    kernel_ms = float(per.mean())
    achieved = step_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] mode={args.mode} {args.steps} steps: {ms_step:.4f} ms/step (events), "
        f"launches first {per[0]:.4f} min {per.min():.4f} median {np.median(per):.4f} max {per.max():.4f} ms, "
        f"B_step={step_bytes / 1e9:.3f} GB -> {achieved:.0f} GB/s")

    q_gpu = qo.numpy()
    f_gpu = fo.numpy().astype(np.float64)
    result = {
        "metric": "reduced steps/s (decode+project+unproject), r=200 @128x192x128; % HBM peak",
        "value": round(1000.0 / ms_step, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"fused decode+project+unproject, grid {grid[0]}x{grid[1]}x{grid[2]}, r={r}, "
                        f"percent={percent}, nBits=32",
            "grid": list(grid), "r": r, "blocks_per_component": basis.nB,
            "tokens": list(tokens), "bytes_per_step": step_bytes,
            "accumulation": "fp32 per segment dot, fp64 per wave and across workgroups",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic_from_profiles(args.workload),
        },
    }
    result["launch_ms"] = {"first": round(float(per[0]), 5), "min": round(float(per.min()), 5),
                           "median": round(float(np.median(per)), 5), "max": round(float(per.max()), 5),
                           "mean": round(kernel_ms, 5)}
    if halves:
        result["halves"] = halves
    result["settle_launches"] = settle
    mops = mfma_mops_from_profiles(args.workload)
    if mops is not None and args.mode == "step":
        # the inverse DCT is the only MFMA work (BF16 hi + lo products)
        tflops = mops * 512 / (kernel_ms * 1e-3) / 1e12
        result["mfma"] = {"flop_per_launch": mops * 512, "achieved_tflops": round(tflops, 2),
                          "peak_tflops": 2500.0, "frac": round(tflops / 2500.0, 5),
                          "counter": "SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 from the committed PMC profile"}

    if spatial is not None:
        result["spatial_decode"] = spatial

    if not args.no_cpu:
        ns = min(args.cpu_sample_cols, r)
        sps, t_full, t_sample, q_s, f_s = cpu_baseline(bufs, v, q, r, ns)
        result["cpu_baseline"] = {
            "value": round(sps, 5), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 oracle project+unproject of all {r} columns, one step ({t_sample:.1f} s)" if ns == r else
                       f"FP64 oracle project+unproject of columns 0..{ns - 1} of {r} "
                       f"({t_sample:.1f} s), extrapolated linearly in columns to {t_full:.1f} s/step"),
        }
        # the same FP64 step split over the host's cores (identical results);
        # the reference is single-threaded, so this is a reported extra
        threads = min(args.threads, os.cpu_count() or 1)
        from oracle import oracle as O
        ob = O.Basis(buffers=bufs)
        times = []
        for _ in range(4):  # one warm-up, then the median of three
            t0 = time.perf_counter()
            q_mt = ob.project_mt(v, threads)
            f_mt = ob.unproject_mt(q, np.zeros_like(v), threads)
            times.append(time.perf_counter() - t0)
        t_mt = float(np.median(times[1:]))
        ob.close()
        same = ns == r and np.array_equal(q_mt, q_s) and np.array_equal(f_mt, f_s)
        result["cpu_baseline_all_cores"] = {
            "value": round(1.0 / t_mt, 5), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"the same FP64 step on {threads} threads, median of 3 after one warm-up ({t_mt:.2f} s), "
                      f"results identical: {bool(same)}",
        }
        host = host_cpu()
        result["cpu_baseline"]["host"] = host
        result["cpu_baseline_all_cores"]["host"] = host
        log(f"[bench] CPU oracle step: 1 thread {t_full:.2f} s, {threads} threads {t_mt:.2f} s")
        # parity of the timed GPU step on the sampled columns
        zq = DeviceArray.from_numpy(np.concatenate([q[:ns], np.zeros(r - ns)]))
        fs = DeviceArray.zeros((len(v),), np.float32)
        C.step(basis, basis, dv, qo, zq, fs, stream)
        stream.synchronize()
        f_gpu_s = fs.numpy().astype(np.float64)
        q_gpu_s = qo.numpy()[:ns]
        # the oracle projects the FP64 field; compare against the FP32-rounded input too
        result["parity"] = {
            "q_rel_l2": float(np.linalg.norm(q_gpu_s - q_s) / np.linalg.norm(q_s)),
            "field_rel_l2": float(np.linalg.norm(f_gpu_s - f_s) / np.linalg.norm(f_s)),
            "columns": ns,
        }
        # per-entry error of q relative to its largest entry
        result["parity"]["q_max_entry_rel"] = float(np.max(np.abs(q_gpu_s - q_s)) / np.max(np.abs(q_s)))
        if spatial is not None:
            # every column of the spatial decode against the FP64 oracle's
            # DecodeVectorField (threaded, identical results)
            from oracle import oracle as O
            from cfs_plume_amd.hip import check, hip
            ob = O.Basis(buffers=bufs)
            errs = []
            n = 3 * basis.peeled_cells
            for c0 in range(0, r, 16):
                c1 = min(r, c0 + 16)
                got = np.empty((c1 - c0, n), np.float32)
                check(hip().hipMemcpy(got.ctypes.data, cols.ptr + c0 * n * 4, got.nbytes, 2), "hipMemcpy D2H")
                want = ob.decode_vector_fields_mt(c0, c1, threads)
                errs += [float(np.linalg.norm(got[i] - want[i]) / np.linalg.norm(want[i])) for i in range(c1 - c0)]
            ob.close()
            spatial["column_rel_l2_max"] = max(errs)
            spatial["column_rel_l2_mean"] = float(np.mean(errs))
            spatial["columns_checked"] = len(errs)
            log(f"[bench] decoded columns vs FP64 oracle: max {max(errs):.2e} mean {np.mean(errs):.2e} "
                f"over {len(errs)} columns")
    if args.copy_calibration_mb:
        n = args.copy_calibration_mb << 20
        src = DeviceArray.zeros((n,), np.uint8)
        dst = DeviceArray((n,), np.uint8)
        from cfs_plume_amd.hip import check, hip
        check(hip().hipMemcpy(dst.ptr, src.ptr, n, 3), "hipMemcpy D2D")
        synchronize()
        log(f"[bench] copy calibration: {n} bytes device-to-device")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
