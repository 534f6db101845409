#!/usr/bin/env python3
"""Benchmark: MGPCG projection solves/s on one MI355X.

Workload: the 1024 x 1024 fully liquid box with 64 Gaussian vortices and 16 divergent sources
(seed 2, FP64 state, FP32 PCG vectors with FP64 reductions and FP64 accumulated pressure).  One step is
one complete projection -- rhs formation, coefficient and multigrid-hierarchy rebuild, MG-preconditioned
CG to max|b - A p| <= 1e-6 max|b|, velocity correction and FP64 divergence check -- starting from the
same device-resident input every time (restored from an HBM snapshot, pressure history reset).

Four further legs are appended to the same JSON line: "semi_lagrangian" (config C2's 200-step run with
semi-Lagrangian advection), "mixed_precision" (config C3: the 2048 x 2048
obstacle tank projected with FP32, binary16 and bfloat16 multigrid levels), "batched" (config C4: 256
mixed-size grids projected one after another, then all at once with fs_batch_project) and "serving"
(config C5: a Poisson stream of full simulations served one at a time, then through the serving loop
at each arrival rate).  --sl-steps 0 / --precision-size 0 / --batch-instances 0 / --serving-requests 0 skip
them.

Prints one JSON line.  Only one GPU is used: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "MGPCG solves/s to 1e-6 rel. residual (1024², mixed batch); p50/p99 step latency"
# Algorithmic bytes per fine cell of the finest post-smoothing launch, by the per-level multigrid model
# (c = 4 bytes per FP32 level value): bilinear prolongation 2.25c (read the coarse correction, read and
# write the fine solution) plus nu sweeps of 5c each (read solution, rhs and the two face conductances,
# write solution; a red-black sweep or a Jacobi sweep alike) = 22.25c = 89 B at nu = 4.  The kernels fuse
# all passes through LDS, so their HBM traffic (measured with rocprofv3 counters) is far lower than this.
def smooth_bytes_per_cell(nu=4):
    return (2.25 + 5.0 * nu) * 4.0


SMOOTH_BYTES_PER_CELL = smooth_bytes_per_cell(4)
# Tile variants (tx, ty, threads) of the fused smoothers, as fs_mg_level_tiles numbers them.
RB_VARIANTS = {0: (64, 32, 256), 1: (32, 32, 256), 2: (32, 16, 256), 3: (16, 16, 256), 4: (64, 32, 512),
               5: (32, 32, 512), 6: (32, 32, 384), 7: (64, 32, 640), 8: (32, 64, 640), 9: (64, 32, 960),
               10: (32, 64, 960)}
JACOBI_VARIANTS = {0: (64, 32, 512), 1: (32, 32, 256), 2: (32, 16, 256), 3: (16, 16, 256), 4: (32, 32, 512),
                   5: (64, 64, 1024), 6: (64, 32, 1024), 7: (32, 64, 1024)}
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(HERE, "profiles")
TRACE_PROFILE = os.path.join(HERE, "profiles", "bench_1024_kernel_trace.md")
SQ_PROFILE = os.path.join(HERE, "profiles", "k_rb_multi_sq_counters.md")
TAIL_KERNEL = 6   # fs_time_kernel id of the single-workgroup coarse tail
# Algorithmic bytes per fine cell of one whole projection, by the per-iteration model: 15 FP32 CG vector
# passes (operator + dot 4, update + max 6, dot + direction 5; s = 4 B) plus the V(nu,nu) cycle's
# (10 nu + 5.5) c per level over a hierarchy summing to 4/3 of the fine level (c = 4 B): at nu = 4,
# 60 + 242.7 = 302.7 B per PCG iteration and per V-cycle; one extra V-cycle, rz and direction start
# the solve, and set-up, rhs, correction and the FP64 divergence check add about 65 B.
def solve_bytes_per_cell_iteration(nu=4):
    return round(60.0 + (10.0 * nu + 5.5) * 4.0 * 4.0 / 3.0, 1)


SOLVE_BYTES_PER_CELL_ITERATION = solve_bytes_per_cell_iteration(4)
SOLVE_SETUP_BYTES_PER_CELL = 65.0


def measured_traffic(kernel, size):
    """HBM bytes per launch of the fine post-smoother from a committed rocprofv3 counter summary
    (profiles/*_traffic.md: FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size) whose title names
    this kernel and grid size, or None when none does (the autotuner chose another tile or smoother)."""
    try:
        names = sorted(n for n in os.listdir(PROFILES) if n.endswith("_traffic.md"))
    except OSError:
        return None
    for name in names:
        with open(os.path.join(PROFILES, name)) as f:
            lines = f.read().splitlines()
        if not lines or f"`{kernel}`" not in lines[0] or f"{size} x {size}" not in lines[0]:
            continue
        for line in lines:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def traced_kernels():
    """(kernel, share of traced kernel time) rows of the committed kernel-trace summary of this benchmark,
    largest first, or [] when there is none."""
    try:
        with open(TRACE_PROFILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return []
    rows = []
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].startswith("`"):
            try:
                rows.append((cells[0].strip("`"), float(cells[4]) / 100.0))
            except ValueError:
                pass
    return rows


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(wl, tau, budget_s, smoother=None):
    """The FP64 CPU oracle (the restatement of UniformCPU, with its red-black V(4,4) cycle) on the same
    projection, timed as the reference runs it: its passes on min(8, host cores) workers
    (FluidGridCPU.cpp:25-31, PressureMultigrid.cpp:13-21), and again on one thread.  Median of the timed
    projections after one warm-up, half the budget each.  Returns (the oracle's projected state, the baseline,
    the oracle's projected state with `smoother` = (kind, sweeps, coarse sweeps), the GPU's choice, or
    None)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O  # noqa: E402  (checker / baseline only)

    g = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    g.configure(wl.liquid_density, wl.gas_density)
    g.set_gravity(wl.gravity_y)

    def timed(threads, budget):
        O.set_threads(threads)
        times, iters = [], []
        t_start = time.perf_counter()
        for k in range(201):
            g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
            t0 = time.perf_counter()
            r = g.project(wl.dt, 200, tau)
            if k > 0:   # the first projection is the warm-up
                times.append(time.perf_counter() - t0)
                iters.append(r["iterations"])
            if len(times) >= 5 and time.perf_counter() - t_start >= budget:
                break
        O.set_threads(1)
        return float(np.median(times)), len(times), int(np.median(iters))

    cores = host_cores()
    workers = min(8, cores)
    med_n, count_n, it_n = timed(workers, 0.5 * budget_s)
    med_1, count_1, it_1 = timed(1, 0.5 * budget_s)
    state = g.download()   # the oracle's projection of the input, for the GPU result's error
    same = None
    if smoother is not None:
        O.set_threads(min(8, cores))
        O.set_smoother(*smoother)
        try:
            g.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
            g.project(wl.dt, 200, tau)
            same = g.download()
        finally:
            O.set_smoother(O.SMOOTH_RBGS, 4, 4)
            O.set_threads(1)
    return state, same, {
        "value": 1.0 / med_n,
        "unit": "solves/s",
        "cores": workers,
        "kind": "port",
        "sample": f"{count_n} FP64 projections of the same {wl.width}x{wl.height} input on {workers} threads "
                  f"(the reference's worker count min(8, cores); median {med_n * 1e3:.1f} ms, {it_n} PCG iterations) "
                  f"on {cpu_model()}, {cores} cores available",
        "single_thread": {"value": 1.0 / med_1, "unit": "solves/s", "cores": 1,
                          "sample": f"{count_1} projections, median {med_1 * 1e3:.1f} ms, {it_1} PCG iterations"},
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
    }


def parse_smoother(text):
    """'red_black:N:C' / 'chebyshev:N:C' -> (kind, N, C); empty -> None."""
    if not text:
        return None
    name, *counts = text.split(":")
    return (1 if name == "chebyshev" else 0, int(counts[0]) if counts else 4, int(counts[1]) if len(counts) > 1 else 4)


def oracle_error(gpu, oracle):
    """Relative L2 distance of a GPU projection from the FP64 oracle's projection of the same input to the
    same tolerance: u, v, and the pressure with each field's mean removed (the pressure is defined up to a
    constant)."""
    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    pg = gpu.pressure - gpu.pressure.mean()
    po = oracle["pressure"] - oracle["pressure"].mean()
    return {"u": rel(gpu.u, oracle["u"]), "v": rel(gpu.v, oracle["v"]), "pressure": rel(pg, po),
            "against": "FP64 CPU oracle, same input and tolerance"}


def dam_break_leg(steps=100, seed=1):
    """Config C1: the 128 x 128 dam break, `steps` full steps (reference transport, gravity, projection to
    |div u| <= 1e-6), on the GPU and on the FP64 CPU oracle (16384 cells: the reference runs every pass
    of this size on one worker), each timed end to end from the same initial state."""
    from fluidsim_amd import workloads as WL
    from fluidsim_amd import FluidGridHIP

    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O  # noqa: E402  (baseline only)

    wl = WL.dam_break(128, 128, seed=seed)
    g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    WL.load_into(g, wl)
    g.step(wl.dt, 200, 1e-6)   # warm-up: hierarchy, graphs, tile choice
    WL.load_into(g, wl)
    iters = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        r = g.step(wl.dt, 200, 1e-6)
        if not r.converged:
            raise RuntimeError("C1 GPU step did not converge")
        iters += r.iterations
    gpu_s = time.perf_counter() - t0
    g.destroy()
    o = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    o.configure(wl.liquid_density, wl.gas_density)
    o.set_gravity(wl.gravity_y)
    o.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    O.set_threads(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(wl.dt, 200, 1e-6)
    cpu_s = time.perf_counter() - t0
    return {
        "workload": f"dam_break 128x128, {steps} steps (PLIC transport, gravity, MG-PCG to |div u| <= 1e-6), "
                    f"seed {seed}",
        "gpu_steps_per_s": steps / gpu_s,
        "gpu_pcg_iterations_mean": iters / steps,
        "cpu_steps_per_s": steps / cpu_s,
        "cpu_kind": "port (FP64 oracle, one thread)",
    }


# fs_mg_precision values of --serving-precision (None: the leg's own choice)
SERVING_PRECISIONS = {"default": None, "fp32": 0, "fp16": 1, "bf16": 2}


def serving_leg(count, rates, steps, lanes, seed=5, max_size=2048, repeats=3, batch_cells=0,
                compare_batch_cells=0, smoother=None, precision=None, viscosity_sweeps=0):
    """Config C5 on one GPU: a seeded Poisson stream of `count` simulations (sides log-uniform in
    [128, max_size], transport + viscosity + surface tension + MG-PCG, `steps` resident steps each).
    It is first served one request at a time on one stream with FP32 multigrid levels (serial plain), then
    fed into the serving loop (`lanes` concurrent streams, HBM-budget admission) at each arrival rate in
    `rates` (per second): requests of at most `batch_cells` cells step together in batched launches with
    FP32 levels (the batched solves' format), larger ones one by one with bfloat16 multigrid conductances.
    Reports aggregate steps/s, per-step latency p50/p99 and queueing delay.  The leg varies from run to run
    (which large grids step side by side depends on arrival and thread timing), so the serial pass and each
    rate run `repeats` times and the median run is reported, with every run's rate listed.  With
    `compare_batch_cells`, one more run at the last rate batches the requests of at most that many cells
    (fs_batch_step: one workgroup per instance in the persistent solves).  `smoother` (kind, sweeps, coarse
    sweeps) is the pressure smoother of the served FP32-level requests (one more run at the last rate then
    serves them with the reference's cycle and bfloat16 conductances); the serial pass keeps the
    reference's V(4,4) cycle.  `precision` (an fs_mg_precision) overrides the served requests' multigrid storage.
    `viscosity_sweeps` (fs_set_viscosity_smoother) is the viscosity smoother of the same requests."""
    from fluidsim_amd import FS_MG_BF16, FS_MG_FP32
    from fluidsim_amd import workloads as WL
    from fluidsim_amd.serving import Request, Server

    t_start = time.perf_counter()

    def note(msg):
        print(f"[serving leg {time.perf_counter() - t_start:7.2f} s] {msg}", file=sys.stderr, flush=True)

    def stream_at(rate, precision=None, cells=batch_cells, choice=None):
        out = []
        for t, kw in WL.serving_stream(seed, count, 128, max_size, rate, steps):
            small = kw["width"] * kw["height"] <= cells
            p = precision if precision is not None else (FS_MG_FP32 if small or choice else FS_MG_BF16)
            extra = {}
            if choice and not small:
                extra = dict(pressure_smoother=choice[0], smoother_sweeps=choice[1], coarse_sweeps=choice[2],
                             viscosity_sweeps=viscosity_sweeps)
            out.append((t, Request(multigrid_precision=p, **kw, **extra)))
        return out

    base = stream_at(rates[0], FS_MG_FP32)
    total_steps = sum(r.steps for _, r in base)
    serial_runs = []
    for _ in range(repeats):
        with Server(lanes=1) as srv:
            t0 = time.perf_counter()
            for _, r in base:
                rep = srv.wait(srv.submit(r))
                if not rep.done:
                    raise RuntimeError(f"serial request failed: {rep.error}")
            serial_runs.append((time.perf_counter() - t0, srv.stats()))
    serial_runs.sort(key=lambda v: v[0])
    serial_s, serial = serial_runs[len(serial_runs) // 2]
    note(f"served one at a time: {total_steps / serial_s:.1f} steps/s")
    # one more serial pass with the served requests' smoothers, to separate what the smoothers give from
    # what serving several requests at once gives
    serial_same = None
    if smoother:
        with Server(lanes=1) as srv:
            t0 = time.perf_counter()
            for _, r in stream_at(rates[0], FS_MG_FP32, choice=smoother):
                rep = srv.wait(srv.submit(r))
                if not rep.done:
                    raise RuntimeError(f"serial request failed: {rep.error}")
            serial_same = time.perf_counter() - t0
        note(f"served one at a time with the served smoothers: {total_steps / serial_same:.1f} steps/s")

    def serve(stream, cells):
        with Server(lanes=lanes, batch_cells=cells) as srv:
            ids = []
            t0 = time.perf_counter()
            for t, r in stream:
                delay = t0 + t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit(r))
            srv.drain()
            served_s = time.perf_counter() - t0
            st = srv.stats()
            reps = [srv.wait(i) for i in ids]
        failed = [r.error for r in reps if not r.done]
        if failed:
            raise RuntimeError(f"served request failed: {failed[0]}")
        return (served_s, st, [r.queue_ms for r in reps], sum(r.setup_ms for r in reps),
                sum(r.teardown_ms for r in reps))

    runs = []
    for rate in rates:
        stream = stream_at(rate, precision=precision, choice=smoother)
        samples = []
        for _ in range(repeats):
            samples.append(serve(stream, batch_cells))
        samples.sort(key=lambda v: v[0])
        served_s, st, queue, setup_ms, teardown_ms = samples[len(samples) // 2]
        note(f"rate {rate}/s: {total_steps / served_s:.1f} steps/s")
        runs.append({
            "rate_per_s": rate,
            "arrival_span_s": stream[-1][0],
            "steps_per_s": total_steps / served_s,
            "steps_per_s_runs": [total_steps / v[0] for v in samples],
            "speedup_vs_serial": serial_s / served_s,
            "step_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]},
            "queue_ms": {"p50": float(np.percentile(queue, 50)), "p99": float(np.percentile(queue, 99))},
            "peak_resident_bytes": st["peak_resident_bytes"],
            "lane_ms_creating_instances": setup_ms,
            "lane_ms_retiring_instances": teardown_ms,
        })
    # the last rate once more with the reference's V(4,4) cycle and bfloat16 conductances, when the served
    # requests run another smoother
    reference = None
    if smoother:
        served_s, st = serve(stream_at(rates[-1]), batch_cells)[:2]
        reference = {"rate_per_s": rates[-1], "smoother": "the reference's red-black V(4,4), bfloat16 conductances",
                     "steps_per_s": total_steps / served_s, "speedup_vs_serial": serial_s / served_s,
                     "step_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]}}
    compare = None
    if compare_batch_cells > 0:
        served_s, st = serve(stream_at(rates[-1], precision=precision, cells=compare_batch_cells, choice=smoother),
                             compare_batch_cells)[:2]
        compare = {"batch_cells": compare_batch_cells, "rate_per_s": rates[-1],
                   "batched_requests": sum(r.width * r.height <= compare_batch_cells for _, r in base),
                   "steps_per_s": total_steps / served_s, "speedup_vs_serial": serial_s / served_s,
                   "step_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]}}
    return {
        "workload": f"{count} requests, sides log-uniform 128..{max_size}, {steps} steps each (PLIC transport, "
                    "viscosity nu 1e-3..1, surface tension 0.01..0.07, MG-PCG to a 1e-6 relative residual "
                    "max|b - A p| <= 1e-6 max|b| in every step; FP32 "
                    "multigrid levels served one at a time; in the serving loop, requests of at most "
                    f"{batch_cells} cells batched with FP32 levels, larger ones "
                    + (f"with FP32 levels and the smoother {smoother}" if smoother else "with bfloat16 conductances")
                    + "), "
                    f"Poisson arrivals, seed {seed}",
        "lanes": lanes,
        "batch_cells": batch_cells,
        "served_smoother": list(smoother) if smoother else None,
        "served_viscosity_sweeps": viscosity_sweeps if smoother else 0,
        "batched_requests": sum(r.width * r.height <= batch_cells for _, r in base),
        "cells_mean": float(np.mean([r.width * r.height for _, r in base])),
        "steps": total_steps,
        "serial_steps_per_s": total_steps / serial_s,
        "serial_steps_per_s_runs": [total_steps / v[0] for v in serial_runs],
        "repeats": repeats,
        "serial_step_ms": {"p50": serial["step_p50_ms"], "p99": serial["step_p99_ms"]},
        "served": runs,
        "served_reference_cycle": reference,
        "serial_with_served_smoothers": None if serial_same is None else {
            "steps_per_s": total_steps / serial_same,
            "served_speedup_over_it": [serial_same / (total_steps / r["steps_per_s"]) for r in runs]},
        "served_batched": compare,
    }


def semi_lagrangian_leg(size=1024, steps=200, seed=2, smoother=None):
    """Config C2's second part: `steps` full steps (semi-Lagrangian advection, gravity, projection) of the
    vortex box, timed end to end, twice: every projection to a 1e-6 relative residual of its own rhs
    (fs_set_relative_tolerance: max|b - A p| <= 1e-6 max|b| in every step, C2's stopping rule), and with the
    tolerance fixed at 1e-6 of the initial max divergence (after the first steps the warm-started pressure
    nearly meets it, so that run is mostly advection).  `smoother` = (kind, sweeps, coarse sweeps) or None."""
    from fluidsim_amd import FS_TRANSPORT_SEMI_LAGRANGIAN, FluidGridHIP
    from fluidsim_amd import workloads as WL

    wl = WL.vortex_box(size, size, seed=seed)
    out = {"workload": f"vortex_box {size}x{size}, {steps} steps of semi-Lagrangian advection + gravity + "
                       "projection, seed 2"}
    for relative in (True, False):
        g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
        WL.load_into(g, wl)
        g.setTransport(FS_TRANSPORT_SEMI_LAGRANGIAN)
        if smoother:
            g.setPressureSmoother(*smoother)
        tau = 1e-6 * g.maxDivergence()
        if relative:
            g.setRelativeTolerance(True)
            tol = 1e-6
        else:
            tol = tau
        g.snapshotSave()
        g.step(wl.dt, 200, tol)   # warm-up: hierarchy, graphs
        g.snapshotRestore()
        its = []
        t0 = time.perf_counter()
        for _ in range(steps):
            r = g.step(wl.dt, 200, tol)
            if not r.accepted():
                raise RuntimeError(f"semi-Lagrangian step rejected: {r}")
            its.append(r.iterations)
        elapsed = time.perf_counter() - t0
        div = g.maxDivergence()
        g.destroy()
        entry = {"steps_per_s": steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
                 "pcg_iterations_mean": float(np.mean(its))}
        if relative:
            entry["stopping_rule"] = "1e-6 relative residual in every step"
            out["relative_residual"] = entry
        else:
            entry["stopping_rule"] = "1e-6 of the initial max divergence"
            entry["final_max_divergence_over_tol"] = div / tau
            out["initial_divergence"] = entry
    # the headline figure of the leg: C2's own rule
    out["steps_per_s"] = out["relative_residual"]["steps_per_s"]
    out["pcg_iterations_mean"] = out["relative_residual"]["pcg_iterations_mean"]
    return out


def smoother_model_bytes(coef_bytes, coarse_bytes, nu=4):
    """Algorithmic bytes per fine cell of the finest post-smoothing launch: nu sweeps of (read solution
    and rhs, FP32 on the finest level, read two conductances, write the solution) plus the bilinear
    prolongation (read a quarter coarse value, read and write the fine solution)."""
    return nu * (12 + 2 * coef_bytes) + 8 + 0.25 * coarse_bytes


def precision_leg(size=2048, seed=3, steps=3, kernel_reps=50, stall_budget=400, only=None):
    """Config C3 on one GPU: the obstacle tank (40 seeded circles and boxes over 15% of the cells, liquid
    in the bottom 70%, vortex velocity) projected to 1e-6 of its initial max divergence with each
    multigrid storage mode, from the same device-resident input every solve.  Reports solves/s, PCG
    iterations, the finest post-smoother's launch time and model bytes, each result's distance from the
    FP32 result, and each result's error against the FP64 CPU oracle's projection of the same input to the
    same tolerance (relative L2 of u, v and of the pressure over the fluid cells, each pressure taken
    relative to its fluid-cell mean; the oracle runs on min(8, host cores) workers).  The bfloat16-vector
    mode gets `stall_budget` iterations."""
    from fluidsim_amd import FS_MG_BF16, FS_MG_BF16_VECTORS, FS_MG_FP16, FS_MG_FP32, FluidGridHIP
    from fluidsim_amd import workloads as WL

    t_start = time.perf_counter()

    def note(msg):
        print(f"[precision leg {time.perf_counter() - t_start:7.2f} s] {msg}", file=sys.stderr, flush=True)

    wl = WL.obstacle_tank(size, size, seed=seed)
    mask = wl.meta["solid"]
    note(f"generated the {size}^2 obstacle tank ({mask.mean():.3f} solid)")
    # (name, storage, conductance bytes, coarse-vector bytes); fp32_tuned adds the smoother the tuner picks
    modes = [("fp32", FS_MG_FP32, 4, 4), ("fp32_tuned", FS_MG_FP32, 4, 4), ("fp16", FS_MG_FP16, 2, 4),
             ("bf16", FS_MG_BF16, 2, 4), ("bf16_vectors", FS_MG_BF16_VECTORS, 2, 2)]
    if only:
        modes = [m for m in modes if m[0] in only]
    out = {"workload": f"obstacle_tank {size}x{size}: 40 circles/boxes over {mask.mean():.3f} of the cells, "
                       f"liquid in the bottom 70% (1000:1), vortex velocity, seed {seed}; one projection per "
                       "solve to 1e-6 of the initial max divergence", "modes": {}}
    # the FP64 oracle's projection of the same input (checker only)
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O  # noqa: E402

    fluid = mask.ravel() == 0
    o = O.OracleGrid(wl.width, wl.height, wl.cell_size)
    o.configure(wl.liquid_density, wl.gas_density)
    o.set_gravity(wl.gravity_y)
    o.set_solid(mask)
    o.upload(wl.fraction, wl.u, wl.v, None, wl.liquid_mass, wl.gas_mass)
    tau_o = 1e-6 * o.max_divergence()
    O.set_threads(min(8, host_cores()))
    t_o = time.perf_counter()
    ro = o.project(wl.dt, 2000, tau_o)
    t_o = time.perf_counter() - t_o
    O.set_threads(1)
    so = o.download()
    po = so["pressure"][fluid] - so["pressure"][fluid].mean()
    out["fp64_oracle"] = {"pcg_iterations": ro["iterations"], "converged": bool(ro["converged"]),
                          "seconds": t_o, "threads": min(8, host_cores())}
    note(f"FP64 oracle: {ro['iterations']} iterations in {t_o:.1f} s")

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    ref_u = None
    for name, mode, coef_bytes, coarse_bytes in modes:
        g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
        WL.load_into(g, wl)
        g.setSolidMask(mask)
        g.setMultigridPrecision(mode)
        tau = 1e-6 * g.maxDivergence()
        budget = stall_budget if mode == FS_MG_BF16_VECTORS else 2000
        tuned = None
        if name == "fp32_tuned":
            kind, nu, coarse, _ = g.tunePressureSmoother(wl.dt, budget, tau)
            tuned = {"smoother": "chebyshev" if kind == 1 else "red_black", "sweeps": nu, "coarse_sweeps": coarse}
            note(f"tuned smoother: {tuned}")
        g.snapshotSave()
        g.snapshotRestore()
        g.project(wl.dt, budget, tau)        # warm-up: hierarchy, graph capture
        results = []
        t0 = time.perf_counter()
        for _ in range(steps):
            g.snapshotRestore()
            results.append(g.project(wl.dt, budget, tau))
        elapsed = time.perf_counter() - t0
        st = g.download()
        u = st.u
        pg = st.pressure[fluid] - st.pressure[fluid].mean()
        post_ms = g.timeKernel(0, kernel_reps)
        model = smoother_model_bytes(coef_bytes, coarse_bytes, tuned["sweeps"] if tuned else 4)
        r = results[-1]
        entry = {
            "solves_per_s": steps / elapsed,
            "ms_per_solve": elapsed / steps * 1e3,
            "pcg_iterations": r.iterations,
            "converged": bool(all(x.converged for x in results)),
            "max_divergence_over_tol": r.max_divergence / tau if tau > 0 else 0.0,
            "fine_post_smoother_us": post_ms * 1e3,
            "fine_post_smoother_model_bytes_per_cell": model,
            "fine_post_smoother_model_gbs": model * size * size / (post_ms * 1e-3) / 1e9,
            "rel_l2_vs_fp64_oracle": {"u": rel(st.u, so["u"]), "v": rel(st.v, so["v"]), "pressure": rel(pg, po)},
        }
        if tuned:
            entry["smoother"] = tuned
        if ref_u is None:
            ref_u = u
        else:
            entry["velocity_rel_l2_vs_fp32"] = float(np.linalg.norm(u - ref_u) / np.linalg.norm(ref_u))
        out["modes"][name] = entry
        note(f"{name}: {entry['ms_per_solve']:.1f} ms/solve, {r.iterations} iterations, "
             f"converged {entry['converged']}, post-smoother {entry['fine_post_smoother_us']:.1f} us")
        g.destroy()
    return out


def batch_leg(count, seed=4, splits=(0,)):
    """Config C4 on one GPU: `count` seeded instances (sides log-uniform 64..512, 30-90% liquid plus a
    drop, vortex velocity, obstacles over 0-25% of the cells), each projected once to 1e-6 of its own initial
    max divergence.  Timed twice
    from the same device-resident inputs: one fs_project per grid in turn (the plain GPU MGPCG), then one
    fs_batch_project over all of them (one persistent launch for all PCG iterations).  Each instance's
    implicit viscosity solve (nu log-uniform 1e-3..1) is timed the same two ways: fs_apply_viscosity one
    grid after another, then fs_batch_apply_viscosity over all of them.  `splits` are the
    fs_context_set_batch_split values timed for the projections (instances above the split run the
    multi-kernel solve beside the persistent launch); the first one is the headline batched number."""
    from fluidsim_amd import (FluidGridHIP, FluidMixture, MixtureComponent, batch_apply_viscosity, batch_project,
                              tune_batch_smoother, FS_TRANSPORT_NONE)
    from fluidsim_amd import workloads as WL

    t_start = time.perf_counter()

    def note(msg):
        print(f"[batch leg {time.perf_counter() - t_start:7.2f} s] {msg}", file=sys.stderr, flush=True)

    wls = WL.batch_instances(seed, count)
    note(f"generated {count} instances")
    grids, taus = [], []
    for wl in wls:
        g = FluidGridHIP(wl.width, wl.height, wl.cell_size)
        nu = wl.meta["nu"]
        g.configureMaterials([], FluidMixture([MixtureComponent("water", 100.0, nu * 100.0)]),
                             FluidMixture([MixtureComponent("air", 0.1, nu * 0.1)]))
        from fluidsim_amd import FluidGridState
        g.upload(FluidGridState(wl.fraction, wl.u, wl.v, np.zeros(0), [wl.liquid_mass], [wl.gas_mass]))
        if wl.meta["solid"] is not None:
            g.setSolidMask(wl.meta["solid"])
        g.setTransport(FS_TRANSPORT_NONE)
        taus.append(1e-6 * g.maxDivergence())
        g.snapshotSave()
        grids.append(g)
        if len(grids) % 64 == 0:
            note(f"created {len(grids)} grids")
    dt = wls[0].dt

    def restore():
        for g in grids:
            g.snapshotRestore()
        grids[0].context.synchronize()

    # warm-up of both paths (graph capture, tile tuning done at creation)
    restore()
    for g, tau in zip(grids, taus):
        g.project(dt, 2000, tau)
    note("serial warm-up done")
    restore()
    batch_project(grids, dt, 2000, taus)
    note("batched warm-up done")
    restore()
    t0 = time.perf_counter()
    serial = [g.project(dt, 2000, tau) for g, tau in zip(grids, taus)]
    serial_s = time.perf_counter() - t0
    note(f"serial {serial_s * 1e3:.1f} ms")
    ctx = grids[0].context
    by_split = {}
    for split in splits:
        ctx.set_batch_split(split)
        restore()
        batch_project(grids, dt, 2000, taus)
        restore()
        t0 = time.perf_counter()
        res = batch_project(grids, dt, 2000, taus)
        by_split[split] = (time.perf_counter() - t0, res)
        note(f"batched (split {split} cells) {by_split[split][0] * 1e3:.1f} ms")
    batch_s, batched = by_split[splits[0]]
    ok = all(r.converged for r in serial) and all(all(r.converged for r in v[1]) for v in by_split.values())
    # the same projections with the red-black sweep counts fs_tune_batch_smoother picks for this batch (at
    # the headline split), batched and one by one with that choice
    ctx.set_batch_split(splits[0])
    restore()
    t0 = time.perf_counter()
    tuned_choice = tune_batch_smoother(grids, dt, 2000, taus)
    tune_s = time.perf_counter() - t0
    note(f"batch tuner: red-black {tuned_choice[0]}/{tuned_choice[1]} in {tune_s:.2f} s")
    restore()
    t0 = time.perf_counter()
    tuned = batch_project(grids, dt, 2000, taus)
    tuned_s = time.perf_counter() - t0
    note(f"batched with the tuned sweeps {tuned_s * 1e3:.1f} ms")
    ctx.set_batch_split(0)
    restore()
    for g, tau in zip(grids, taus):
        g.project(dt, 2000, tau)
    restore()
    t0 = time.perf_counter()
    tuned_serial = [g.project(dt, 2000, tau) for g, tau in zip(grids, taus)]
    tuned_serial_s = time.perf_counter() - t0
    note(f"serial with the tuned sweeps {tuned_serial_s * 1e3:.1f} ms")
    tuned_ok = all(r.converged for r in tuned) and all(r.converged for r in tuned_serial)
    # viscosity: every instance, one by one and then batched (warm-up of both first)
    restore()
    for g in grids:
        g.applyViscosity(dt)
    restore()
    t0 = time.perf_counter()
    vits = [g.applyViscosity(dt) for g in grids]
    visc_s = time.perf_counter() - t0
    note(f"viscosity one by one {visc_s * 1e3:.1f} ms")
    visc_by_split = {}
    for split in splits:
        ctx.set_batch_split(split)
        restore()
        batch_apply_viscosity(grids, dt)
        restore()
        t0 = time.perf_counter()
        res = batch_apply_viscosity(grids, dt)
        ctx.synchronize()
        visc_by_split[split] = (time.perf_counter() - t0, res)
        note(f"viscosity batched (split {split} cells) {visc_by_split[split][0] * 1e3:.1f} ms")
    ctx.set_batch_split(0)
    visc_batch_s, bvits = visc_by_split[splits[0]]
    for k, g in enumerate(grids):
        g.destroy()
        if (k + 1) % 32 == 0:
            note(f"released {k + 1} grids")
    cells = [wl.width * wl.height for wl in wls]
    return {
        "workload": f"{count} instances, sides log-uniform 64..512, liquid 30-90% + drop, vortex velocity, "
                    f"obstacles over 0-25% of the cells, seed {seed}; one projection each to 1e-6 of its "
                    "initial max divergence",
        "masked_instances": int(sum(wl.meta["solid"] is not None for wl in wls)),
        "cells_total": int(sum(cells)),
        "serial_solves_per_s": count / serial_s,
        "batched_solves_per_s": count / batch_s,
        "speedup": serial_s / batch_s,
        "batched_ms": batch_s * 1e3,
        "batch_split_cells": int(splits[0]),
        "batched_ms_by_split": {str(k): v[0] * 1e3 for k, v in by_split.items()},
        "converged_all": bool(ok),
        "pcg_iterations": {"serial_mean": float(np.mean([r.iterations for r in serial])),
                           "batched_mean": float(np.mean([r.iterations for r in batched]))},
        "tuned_sweeps": {"red_black_sweeps": tuned_choice[0], "coarse_sweeps": tuned_choice[1],
                         "tuner_s": tune_s, "batched_solves_per_s": count / tuned_s,
                         "serial_solves_per_s": count / tuned_serial_s, "speedup": tuned_serial_s / tuned_s,
                         "converged_all": bool(tuned_ok),
                         "pcg_iterations_batched_mean": float(np.mean([r.iterations for r in tuned]))},
        "viscosity_solves_per_s_serial": count / visc_s,
        "viscosity_solves_per_s_batched": count / visc_batch_s,
        "viscosity_speedup": visc_s / visc_batch_s,
        "viscosity_batched_ms_by_split": {str(k): v[0] * 1e3 for k, v in visc_by_split.items()},
        "viscosity_iterations": {"serial_mean": float(np.mean(vits)), "batched_mean": float(np.mean(bvits))},
        "projection_plus_viscosity_speedup": (serial_s + visc_s) / (batch_s + visc_batch_s),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--smoother", type=str, default="tuned",
                    help="tuned (fs_tune_pressure_smoother), red_black:N[:C] (N = 1..4) or chebyshev:N[:C] "
                         "(N = 1..5); C = red-black sweeps on the levels of at most 4096 cells (1..4, default 4)")
    ap.add_argument("--fine-tiles", type=str, default="", help="force the finest level's tile variants "
                    "PRE,POST,RESTRICT (profiling runs: the autotuner's choice can shift under a profiler)")
    ap.add_argument("--serving-requests", type=int, default=24, help="size of the C5 serving stream (0: skip)")
    ap.add_argument("--serving-rates", type=str, default="8,64", help="arrival rates to sweep (per second)")
    ap.add_argument("--serving-steps", type=int, default=50)
    ap.add_argument("--serving-lanes", type=int, default=4)
    ap.add_argument("--serving-batch-cells", type=int, default=0,
                    help="C5: requests of at most this many cells step batched (0: every request alone)")
    ap.add_argument("--serving-smoother", type=str, default="chebyshev:8:3",
                    help="C5: pressure smoother of the served requests (FP32 levels), red_black:N:C or "
                         "chebyshev:N:C; empty: the reference's V(4,4) with bfloat16 conductances.  The default, "
                         "8 Chebyshev sweeps and 3 coarse sweeps, is the fastest measured on this stream")
    ap.add_argument("--serving-viscosity-sweeps", type=int, default=4, choices=(0, 2, 4, 6),
                    help="C5: viscosity smoother of the served requests with a pressure smoother choice "
                         "(fs_set_viscosity_smoother: 0 the reference's V(2,2), else Chebyshev sweeps per side; "
                         "the default, 4, is the fastest measured on this stream)")
    ap.add_argument("--serving-compare-batch-cells", type=int, default=65536,
                    help="C5: one extra run at the last rate with this batch size, for comparison (0: skip)")
    ap.add_argument("--serving-repeats", type=int, default=3, help="runs per pass; the median is reported")
    ap.add_argument("--serving-max-size", type=int, default=2048)
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--serving-precision", choices=sorted(SERVING_PRECISIONS), default="default",
                    help="multigrid storage of the served requests (default: bfloat16 conductances above "
                         "the batch size, FP32 below)")
    ap.add_argument("--batch-instances", type=int, default=256, help="size of the C4 batched-solve leg (0: skip)")
    ap.add_argument("--batch-only", action="store_true")
    ap.add_argument("--batch-splits", type=str, default="131072,0",
                    help="fs_context_set_batch_split values to time (cells; the first is reported as batched)")
    ap.add_argument("--precision-size", type=int, default=2048, help="side of the C3 mixed-precision leg (0: skip)")
    ap.add_argument("--precision-steps", type=int, default=3)
    ap.add_argument("--precision-only", action="store_true")
    ap.add_argument("--sl-steps", type=int, default=200, help="C2 semi-Lagrangian steps leg (0: skip)")
    ap.add_argument("--c1-steps", type=int, default=100, help="C1 dam-break steps leg, GPU and CPU (0: skip)")
    ap.add_argument("--precision-modes", type=str, default="", help="subset of fp32,fp32_tuned,fp16,bf16,bf16_vectors")
    args = ap.parse_args()
    splits = tuple(int(v) for v in args.batch_splits.split(",") if v.strip())

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the solver runs on a single GPU; multi-GPU sharding is a design, not code).", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from fluidsim_amd import FS_SMOOTHER_CHEBYSHEV, FS_SMOOTHER_RED_BLACK, FluidGridHIP
    from fluidsim_amd import workloads as WL

    rates = [float(x) for x in args.serving_rates.split(",") if x]
    if args.batch_only:
        print(json.dumps(batch_leg(args.batch_instances or 256, splits=splits)), flush=True)
        return
    if args.precision_only:
        only = [m for m in args.precision_modes.split(",") if m] or None
        print(json.dumps(precision_leg(args.precision_size or 2048, steps=args.precision_steps, only=only)),
              flush=True)
        return
    if args.serving_only:
        print(json.dumps(serving_leg(args.serving_requests or 24, rates, args.serving_steps,
                                     args.serving_lanes, max_size=args.serving_max_size,
                                     batch_cells=args.serving_batch_cells,
                                     compare_batch_cells=args.serving_compare_batch_cells,
                                     smoother=parse_smoother(args.serving_smoother),
                                     precision=SERVING_PRECISIONS[args.serving_precision],
                                     viscosity_sweeps=args.serving_viscosity_sweeps,
                                     repeats=max(1, args.serving_repeats))), flush=True)
        return

    wl = WL.vortex_box(args.size, args.size, seed=2)
    grid = FluidGridHIP(wl.width, wl.height, wl.cell_size)
    WL.load_into(grid, wl)
    tau = 1e-6 * grid.maxDivergence()  # 1e-6 of max |div u*| = 1e-6 relative residual (max-norm)
    # The smoother of the levels above 4096 cells: chosen by fs_tune_pressure_smoother on this input (each
    # candidate projects the same state, which the tuner restores), or fixed on the command line.
    tuned_ms = None
    if args.smoother == "tuned":
        kind, nu, coarse, tuned_ms = grid.tunePressureSmoother(wl.dt, 200, tau)
    else:
        name, *counts = args.smoother.split(":")
        kind = FS_SMOOTHER_CHEBYSHEV if name == "chebyshev" else FS_SMOOTHER_RED_BLACK
        nu = int(counts[0]) if counts else 4
        coarse = int(counts[1]) if len(counts) > 1 else 4
        grid.setPressureSmoother(kind, nu, coarse)
    if args.fine_tiles:
        grid.setMultigridTiles(0, *[int(v) for v in args.fine_tiles.split(",")])
    grid.snapshotSave()

    def one_step():
        grid.snapshotRestore()
        return grid.project(wl.dt, 200, tau)

    for _ in range(args.warmup):
        res = one_step()
    grid.context.synchronize()

    lat = []
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        res = one_step()   # ends with a device synchronisation (result read-back)
        lat.append(time.perf_counter() - s0)
        results.append(res)
    grid.context.synchronize()
    elapsed = time.perf_counter() - t0

    ok = all(r.converged for r in results)
    rel = [r.final_residual / r.rhs_max if r.rhs_max > 0 else 0.0 for r in results]
    iters = [r.iterations for r in results]

    # Dominant kernel: the finest-level post-smoothing launch, timed with HIP events on the grid's stream.
    post_variant = grid.multigridTiles(0)[1]
    chebyshev = kind == FS_SMOOTHER_CHEBYSHEV
    if chebyshev:
        tx, ty, nt = JACOBI_VARIANTS[post_variant]
        fine_kernel = f"k_jacobi_multi<{tx},{ty},1,{nt}>"
        smoother = (f"Chebyshev-weighted damped Jacobi, {nu} sweeps per side, on the levels above 4096 cells "
                    f"(red-black Gauss-Seidel, {coarse} sweeps per side, below)")
        fine_what = f"finest-level prolongation + {nu} Chebyshev-weighted Jacobi sweeps, autotuned tile"
    else:
        tx, ty, nt = RB_VARIANTS[post_variant]
        fine_kernel = f"k_rb_multi<{tx},{ty},4,{1 if post_variant in (0, 4) else 0},1,{nt},float,float,float>"
        smoother = f"red-black Gauss-Seidel, {nu} sweeps per side on the levels above 4096 cells ({coarse} below)"
        fine_what = f"finest-level prolongation + {nu} red-black Gauss-Seidel sweeps, autotuned tile"
    sweep_ms = grid.timeKernel(0, args.kernel_reps)
    tail_ms = grid.timeKernel(TAIL_KERNEL, args.kernel_reps)
    sweep_bytes = smooth_bytes_per_cell(nu) * wl.width * wl.height
    achieved = sweep_bytes / (sweep_ms * 1e-3) / 1e9
    # Reference point: a plain FP32 copy of 64 MiB into another 64 MiB buffer on the same stream.
    copy_ms = grid.timeKernel(10, 20)
    copy_bytes = 2 * 4 * (16 << 20)
    traffic = measured_traffic(fine_kernel, args.size)
    cells = wl.width * wl.height
    per_iteration = solve_bytes_per_cell_iteration(nu)
    solve_bytes = cells * (per_iteration * (float(np.median(iters)) + 1.0) + SOLVE_SETUP_BYTES_PER_CELL)

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": f"vortex_box {wl.width}x{wl.height}: one projection per step to 1e-6 relative "
                        "max-norm residual (64 vortices + 16 sources, seed 2, fully liquid, closed box)",
            "grid": [wl.width, wl.height],
            "preconditioner": f"multigrid V-cycle, FP32 levels: {smoother}",
            "smoother": f"{'chebyshev' if chebyshev else 'red_black'}:{nu}:{coarse}",
            "smoother_choice": ("fs_tune_pressure_smoother on this input (fastest converging of red-black 2-4 and "
                                "Chebyshev 3-5 sweeps, each with 4 or 2 sweeps on the levels of at most 4096 cells: "
                                f"{tuned_ms:.3f} ms per projection)") if tuned_ms else "fixed",
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "pcg_iterations": int(np.median(iters)),
        "converged_all": bool(ok),
        "max_rel_residual": float(max(rel)),
        "roofline": {
            "kernel": f"{fine_kernel} ({fine_what})",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_us": sweep_ms * 1e3,
            "algorithmic_bytes_per_launch": sweep_bytes,
            "copy_kernel_gbs": copy_bytes / (copy_ms * 1e-3) / 1e9,
            "model_note": "algorithmic bytes count every sweep's pass through HBM, as unfused sweeps would make "
                          "it; the fused kernel reads each value once, so frac can exceed 1 (frac_of_measured_"
                          "traffic is the HBM figure)",
        },
        "solve_roofline": {
            "bound": "hbm",
            "algorithmic_bytes_per_solve": solve_bytes,
            "achieved": solve_bytes / (elapsed / args.steps) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": solve_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
            "model": f"{per_iteration} B per fine cell per PCG iteration (and per starting "
                     f"V-cycle) + {SOLVE_SETUP_BYTES_PER_CELL:.0f} B set-up",
        },
    }
    if traffic:
        # the fused smoother moves far fewer bytes than its model: what limits it is instruction issue
        out["roofline"]["frac_of_measured_traffic"] = traffic / (sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
        out["roofline"]["limiter"] = (f"not HBM: measured HBM traffic is {traffic / sweep_bytes:.2f} of the model "
                                      "bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE)")
        try:
            with open(SQ_PROFILE) as f:
                covered = f"`<{fine_kernel.split('<', 1)[1].split(',float')[0]}>`" in f.read().replace(" ", "")
        except OSError:
            covered = False
        if covered:
            out["roofline"]["limiter"] += ("; SQ counters show instruction issue and latency: VALU the busiest pipe, "
                                           "waves waiting a third or more of their time")
            out["roofline"]["limiter_source"] = os.path.relpath(SQ_PROFILE, HERE)
    top = traced_kernels()
    if top:
        dom = {"kernel": top[0][0], "share_of_kernel_time": top[0][1], "source": os.path.relpath(TRACE_PROFILE, HERE)}
        if top[0][0].startswith("k_mg_tail"):
            # the single-workgroup tail: levels from the first untiled one down, V(4,4) at the multigrid
            # model's (10 nu + 5.5) c = 182 B per level cell, plus the FP32 coarse inverse
            shapes = [grid.multigridShape(l) for l in range(grid.multigridLevels())]
            tail = [s for l, s in enumerate(shapes) if grid.multigridTiles(l)[0] < 0]
            last = shapes[-1][0] * shapes[-1][1] - 1
            tail_bytes = 182 * sum(w * h for w, h in tail[:-1]) + 4 * last * last
            tail_us = tail_ms * 1e3
            dom.update({"avg_launch_us": tail_us, "algorithmic_bytes_per_launch": tail_bytes,
                        "achieved_gbs": tail_bytes / (tail_us * 1e-6) / 1e9,
                        "frac": tail_bytes / (tail_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                        "limiter": "latency: one 1024-thread workgroup on one compute unit, about 60 "
                                   "barrier-separated phases per launch"})
        out["roofline"]["dominant_kernel_in_trace"] = dom
    if rank == 0 and not args.no_cpu:
        gpu_state = grid.download()   # the last timed projection
        oracle_state, same_state, out["cpu_baseline"] = cpu_baseline(wl, tau, args.cpu_budget, (int(kind), nu, coarse))
        # against UniformCPU's own cycle (what the reference computes), and against the same smoother
        out["oracle_error"] = oracle_error(gpu_state, oracle_state)
        out["oracle_error"]["against"] = ("FP64 CPU oracle with the reference's red-black V(4,4), same input and "
                                          "tolerance")
        out["oracle_error_same_smoother"] = oracle_error(gpu_state, same_state)
        out["oracle_error_same_smoother"]["against"] = ("FP64 CPU oracle with the GPU's smoother, same input and "
                                                        "tolerance")
    if args.sl_steps > 0:
        out["semi_lagrangian"] = semi_lagrangian_leg(args.size, args.sl_steps, smoother=(int(kind), nu, coarse))
    if args.c1_steps > 0:
        out["dam_break"] = dam_break_leg(args.c1_steps)
    if args.precision_size > 0:
        out["mixed_precision"] = precision_leg(args.precision_size, steps=args.precision_steps)
    if args.batch_instances > 0:
        grid.destroy()
        out["batched"] = batch_leg(args.batch_instances, splits=splits)
    if args.serving_requests > 0:
        grid.destroy()
        out["serving"] = serving_leg(args.serving_requests, rates, args.serving_steps,
                                     args.serving_lanes, max_size=args.serving_max_size,
                                     batch_cells=args.serving_batch_cells,
                                     compare_batch_cells=args.serving_compare_batch_cells,
                                     smoother=parse_smoother(args.serving_smoother),
                                     precision=SERVING_PRECISIONS[args.serving_precision],
                                     viscosity_sweeps=args.serving_viscosity_sweeps,
                                     repeats=max(1, args.serving_repeats))
    if rank == 0:
        print(json.dumps(out), flush=True)
    grid.destroy()


if __name__ == "__main__":
    main()
