#!/usr/bin/env python3
"""Benchmark: FP32 Yee cell-updates/s and fraction of the HBM roofline on one MI355X.

Default workload: the batched taper set.  64 Si tapers (seed 20261203), with lengths
L ~ U[3, 15] um and 8 spline control-point widths ~ U[0.4, 3.2] um.  Each taper is its own grid
of uneven length (170-770 cells along the propagation axis) with a 200 x 100 cross-section of
20 nm cells, rasterised on the device.  Each grid has an 8-cell CPML, a pulsed mode source and an
FP64 overlap monitor with a running DFT at 16 wavelengths 1.50-1.60 um.  All 64 advance in
lockstep through the batched kernels.  This is the largest configuration that fits one GPU:
607.9 M cells.

A "step" is one full time step of every grid: sources, E and H updates with CPML, and monitors.
--steps K times K consecutive steps after --warmup W untimed steps.  The batch first spends a
fixed tuning window picking its launch shape (real steps, before the warmup).

Output: one JSON line.
- value: cell-updates per second over every grid.
- roofline: the dominant kernel, at 48 B per cell-update (6 components x (4 B read + 4 B
  written)).  A split E or H launch advances half a cell-update, so it is credited 24 B per cell.
  Its mean duration comes from a window in which every launch is event-timed.
- traffic: copy-calibrated rocprofv3 counters from profiles/.
- cpu_baseline: the compiled reference (oracle/_ref) timefield3d_apply_timestep_omp_full, FP64
  with Mur faces, on the median-length taper grid of the same set.  It runs in child processes
  with OMP_PROC_BIND=close and OMP_PLACES=cores: once with every physical core of this process's
  CPU set (lscpu), once with the box's OMP_NUM_THREADS share when that differs, and once with
  1 thread.  Each is the best of 3 samples; the fastest leg is `value`.

Other workloads: --workload waveguide | design | surrogate | gradient | shard | dropin.
"""
import argparse
import ctypes as C
import importlib
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Yee cell-updates/s & % HBM roofline; verified taper designs/hour"
HBM_PEAK_GBS = 8000.0
STEP_BYTES = 48  # algorithmic bytes per cell-update: 6 components x (4 B read + 4 B written)
HALF_STEP_BYTES = STEP_BYTES // 2  # one split E or H launch: half a cell-update
SPLIT_KERNEL_BYTES = 36  # what one split launch itself reads and writes (3 comps r+w, 3 read)
TUNE_STEPS = 64  # the batch autotuner's window (11 candidates x 5 steps) on the first run call

WL = importlib.import_module("3d_fdtd_amd.workloads")
DISC = WL.DISC
I, J, K = WL.WG_I, WL.WG_J, WL.WG_K
NPML = WL.WG_NPML
N_RUN = 4000


def bench_tapers(fd, args):
    """Batched taper workload: whole-job cell-updates/s over every grid (one JSON line)."""
    W, S = args.warmup, args.steps
    KW = 10  # kernel-timing window (every E / H launch event-timed, no graph)
    print(f"bench: building {args.tapers} taper grids (device raster)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    grids, dt, omegas, lengths, widths, _, _, _ = WL.tapers(args.tapers, TUNE_STEPS + W + S + KW)
    b = fd.Batch(grids)
    if args.kchunk:
        b.set_shape(args.kchunk, args.kchunk <= 4, True)
    setup_s = time.perf_counter() - t0
    print(f"bench: set-up {setup_s:.1f} s; tuning window {TUNE_STEPS} steps", file=sys.stderr, flush=True)
    b.run(dt, 0, TUNE_STEPS)
    info = b.info()
    print(f"bench: shape {info['shape']}; warmup {W} steps", file=sys.stderr, flush=True)
    n0 = TUNE_STEPS
    b.run(dt, n0, W)
    n0 += W
    print(f"bench: timing {S} steps", file=sys.stderr, flush=True)
    L = fd.lib()
    fd.dev_check(L.fdtd_dev_batch_timer_start(b.b), "timer")
    t0 = time.perf_counter()
    b.run(dt, n0, S)
    ms = C.c_double()
    fd.dev_check(L.fdtd_dev_batch_timer_stop(b.b, C.byref(ms)), "timer")
    wall = time.perf_counter() - t0
    n0 += S
    elapsed_s = max(ms.value / 1e3, wall)
    # kernel durations: a separate window in which every launch carries its own event pair
    # (graph replays carry none)
    fd.dev_check(L.fdtd_dev_batch_kernel_timing(b.b, 1), "kernel timing")
    b.run(dt, n0, KW)
    cnt, kms = (C.c_int * 2)(), (C.c_double * 2)()
    fd.dev_check(L.fdtd_dev_batch_kernel_times(b.b, cnt, kms), "kernel times")
    fd.dev_check(L.fdtd_dev_batch_kernel_timing(b.b, 0), "kernel timing off")
    e_us, h_us = (kms[w] / cnt[w] * 1e3 if cnt[w] else 0.0 for w in range(2))
    cells = info["cells"]
    value = cells * S / elapsed_s
    fused = info.get("fused", False)
    if fused:  # one launch per step: a whole cell-update per cell
        dom, dom_name, per_launch = e_us, "batch_step_fused", STEP_BYTES * cells
    else:
        dom = max(e_us, h_us)
        dom_name = "batch_update_E" if e_us >= h_us else "batch_update_H"
        per_launch = HALF_STEP_BYTES * cells
    achieved = per_launch / (dom * 1e-6) / 1e9 if dom > 0 else 0.0
    traffic = read_traffic("tapers_batch_traffic.json") if args.tapers == 64 else None
    t_mon = [grids[q].overlap(grids[q].mon, len(omegas))[0] for q in (0, len(grids) - 1)]
    print(f"bench: monitor |a_out| at 1.55 um (grids 0, last): {abs(t_mon[0][7]):.3e} {abs(t_mon[1][7]):.3e}",
          file=sys.stderr, flush=True)
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": S, "warmup": W,
           "ms_per_step": elapsed_s * 1e3 / S, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batched tapers: {args.tapers} Si tapers (seed {WL.TAPER_SEED}), lengths "
                                  "U[3,15] um, 8 spline control-point widths U[0.4,3.2] um, 200x100 cross-section "
                                  "of 20 nm cells, 8-cell CPML, mode source, FP32 fields, FP64 overlap monitor x 16 "
                                  "wavelengths per taper; all grids advance in lockstep (batched E/H kernels "
                                  "over a per-grid descriptor table, launch shape autotuned, hipGraph-captured "
                                  "steps)",
                      "grids": args.tapers, "cells_total": cells, "setup_s": round(setup_s, 2),
                      "tuning_steps_before_warmup": TUNE_STEPS,
                      "lengths_cells": [WL.taper_length_cells(x) for x in lengths]},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic.get(dom_name) if traffic else None,
                        "kernel": dom_name, "kernel_avg_us": round(dom, 1), "e_us": round(e_us, 1),
                        "h_us": round(h_us, 1), "timed_launches": int(max(cnt[0], cnt[1])),
                        "kernel_timing_window_steps": KW,
                        "algorithmic_bytes_per_launch": per_launch,
                        "bytes_per_cell_update": STEP_BYTES,
                        "frac_of_value_48B": round(STEP_BYTES * value / 1e9 / HBM_PEAK_GBS, 4),
                        "autotuned_shape": info["shape"], "tuned_ms_per_step": info["tuned"]}}
    b.free()
    for g in grids:
        g.free()
    if not args.no_cpu:
        print("bench: CPU baseline (compiled reference, child processes)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_taper(args, lengths)
    print(json.dumps(out), flush=True)


def physical_cores():
    """(physical cores, logical CPUs) in this process's CPU set: lscpu's distinct (socket, core)
    pairs over the CPUs of the affinity mask."""
    aff = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return len(aff), len(aff)
    cores = set()
    for ln in out.splitlines():
        if not ln or ln.startswith("#"):
            continue
        cpu, core, sock = ln.split(",")[:3]
        if int(cpu) in aff:
            cores.add((sock, core))
    return (len(cores) or len(aff)), len(aff)


def cpu_baseline_taper(args, lengths):
    """The compiled reference's full step on the median-length taper grid, in child processes
    (OpenMP reads its environment when the runtime starts)."""
    q = int(np.argsort(lengths)[len(lengths) // 2])
    ncores, ncpu = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    legs = [ncores] + ([share] if 0 < share < ncores else []) + [1]
    results = []
    for n in legs:
        env = dict(os.environ, OMP_NUM_THREADS=str(n), OMP_PROC_BIND="close", OMP_PLACES="cores")
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", "--cpu-taper", str(q), "--tapers",
               str(args.tapers), "--cpu-sample", "3.0" if n > 1 else "1.0"]
        try:
            r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
        except subprocess.TimeoutExpired:
            print(f"bench: CPU leg with {n} threads timed out", file=sys.stderr, flush=True)
            continue
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode != 0 or not line:
            print(f"bench: CPU leg with {n} threads failed: {r.stderr[-400:]}", file=sys.stderr, flush=True)
            continue
        res = json.loads(line[-1])
        res["threads"] = n
        results.append(res)
        print(f"bench: CPU leg {n} threads: {res['value'] / 1e6:.1f} M cell-updates/s", file=sys.stderr, flush=True)
    if not results:
        return None
    best = max(results, key=lambda r: r["value"])
    return {"value": best["value"], "unit": "cell-updates/s", "cores": best["threads"], "kind": best["kind"],
            "sample": f"{best['kind']} FP64 full step (Mur faces) on taper {q} of the set "
                      f"({best['grid']}), best of 3 samples of {best['steps']} steps each; "
                      f"OMP_PROC_BIND=close OMP_PLACES=cores; host: {ncores} physical cores / {ncpu} CPUs "
                      "in the CPU set",
            "legs": [{"threads": r["threads"], "value": r["value"], "best_s_per_step": r["best_s_per_step"]}
                     for r in results]}


def cpu_leg(args):
    """Child process of cpu_baseline_taper: time the compiled reference's full step (the
    clean-room FP64 oracle where the reference library is absent) on taper --cpu-taper of the
    set; prints one JSON line."""
    from oracle import ref_py

    lengths, widths = WL.taper_geometry(args.tapers)
    q = args.cpu_taper
    ep, K = WL.taper_permittivity(lengths[q], widths[q])
    I, J = WL.TAPER_I, WL.TAPER_J
    dt = 0.99 * DISC / math.sqrt(3)
    rng = np.random.default_rng(0)
    if ref_py.available():
        R = ref_py.lib()
        tf = R.timefield3d_new(I, J, K, DISC)
        A = ref_py.tf_arrays(tf)
        A["ep"][:] = ep
        A["fields"][:] = rng.uniform(-1, 1, A["fields"].size)

        def step():
            R.timefield3d_apply_timestep_omp_full(tf, dt)

        kind = "reference"
    else:
        from oracle import oracle_py as O

        o = O.Grid(I, J, K, DISC)
        o.ep[:] = ep
        o.f[:] = rng.uniform(-1, 1, o.f.size)

        def step():
            o.step(dt)

        kind = "port"
    del ep
    step()  # warm
    t0 = time.perf_counter()
    step()
    est = time.perf_counter() - t0
    n = max(1, int(args.cpu_sample / max(est, 1e-6)))
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(n):
            step()
        best = min(best, (time.perf_counter() - t0) / n)
    cells = I * J * K
    print(json.dumps({"value": cells / best, "kind": kind, "steps": n, "best_s_per_step": best,
                      "grid": f"{I}x{J}x{K} library cells"}), flush=True)


DESIGN_DATA = os.path.join(ROOT, "3d_fdtd_amd", "data")


def design_family(design):
    """The taper family of the design loop (that of the batched 64-taper set): 20 nm cells, 200 x 100
    cross-section, 8-cell CPML, lengths U[3, 15] um, 8 spline control widths U[0.4, 3.2] um
    (fill-fraction raster), 6000 steps, 16 wavelengths 1.50-1.60 um, TE0 port modes from the mode
    solver."""
    return design.TaperFamily(disc=DISC, I=WL.TAPER_I, J=WL.TAPER_J, core_cells=WL.TAPER_CORE, npml=WL.TAPER_NPML,
                              port=0, wmin=0.4, wmax=3.2, lmin=3.0, lmax=15.0, steps=6000, lam=(1.50, 1.60),
                              nfreq=16, ports="te0")


def design_data(fd, args):
    """One chunk of the surrogate's FDTD training set: designs drawn from the family with the seed of
    chunk --chunk (of --chunks covering --train designs), simulated in batches of 64; saved to
    DESIGN_DATA/cfg5_train_<chunk>.npz with the device time it took."""
    design = importlib.import_module("3d_fdtd_amd.design")
    fam = design_family(design)
    per = (args.train + args.chunks - 1) // args.chunks
    lo, hi = args.chunk * per, min(args.train, (args.chunk + 1) * per)
    rng = np.random.default_rng(WL.TAPER_SEED + 1000 + args.chunk)
    p = fam.random(rng, hi - lo)
    print(f"bench: design data chunk {args.chunk}: {hi - lo} designs; solving port modes", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    fam.precompute_ports(p)
    t_modes = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = []
    for s0 in range(0, len(p), 64):
        out.append(fam.simulate(p[s0:s0 + 64]))
        print(f"bench: chunk {args.chunk}: {s0 + len(out[-1])} of {len(p)} designs, "
              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    t_sim = time.perf_counter() - t0
    os.makedirs(DESIGN_DATA, exist_ok=True)
    np.savez(os.path.join(DESIGN_DATA, f"cfg5_train_{args.chunk}.npz"), p=p, t=np.concatenate(out), t_sim=t_sim,
             t_modes=t_modes)
    print(json.dumps({"design_data_chunk": args.chunk, "designs": len(p), "fdtd_s": round(t_sim, 1),
                      "mode_solves_s": round(t_modes, 1), "s_per_design": round(t_sim / len(p), 3),
                      "band_T_quantiles": [round(float(v), 4) for v in np.quantile(
                          np.mean(np.abs(np.concatenate(out)) ** 2, axis=1), [0.5, 0.9, 0.95, 0.99])]}), flush=True)


def bench_design(fd, args):
    """Verified taper designs per hour on the batched set's taper family (design_family) at the surrogate
    configuration: an MLP 9 -> 256 x 4 -> 32 (BF16 MFMA) trained on the 4096-design FDTD training set
    (generated beforehand in chunks by --design-phase data), screening 10^5 candidates per
    iteration, the top 64 verified by batched FDTD; against the plain loop (gradient ascent with
    single-geometry forward + adjoint FDTD runs from random starts).

    Acceptance: FDTD band |t|^2 >= spec, spec = the 95th percentile of the training set (random
    designs meet it 1 time in 20).  Strict acceptance adds the surrogate gate: max over the 16
    wavelengths of |T_surrogate - T_FDTD| <= 1e-3 (T = |t|^2) on the accepted design.  The plain loop
    has no surrogate, so its accepted designs count under both.  value = the surrogate loop's
    accepted designs per hour after training (the training set's FDTD time is reported beside).

    The two loops can run in separate calls (--design-iters 0 or --plain-seconds 0): each saves its
    part to DESIGN_DATA, and the line is printed once both parts exist.  --plain-append adds a
    call's plain-loop starts (seeded after the saved ones) to the saved plain part."""
    design = importlib.import_module("3d_fdtd_amd.design")
    fam = design_family(design)
    files = sorted(f for f in os.listdir(DESIGN_DATA) if f.startswith("cfg5_train_")) if os.path.isdir(
        DESIGN_DATA) else []
    if not files:
        raise SystemExit("design workload: no training set in 3d_fdtd_amd/data (run --design-phase data first)")
    chunks = [np.load(os.path.join(DESIGN_DATA, f)) for f in files]
    train_p = np.concatenate([c["p"] for c in chunks])
    train_t = np.concatenate([c["t"] for c in chunks])
    t_data = float(sum(float(c["t_sim"]) for c in chunks))
    T_train = design.band_transmission(train_t)
    spec = float(np.quantile(T_train, 0.95))
    print(f"bench: {len(train_p)} training designs ({t_data:.0f} s of FDTD); spec band T >= {spec:.4f}",
          file=sys.stderr, flush=True)
    # every port width the families can draw, solved once in parallel host processes
    wc = np.arange(fam.port_width_cells(fam.wmin), fam.port_width_cells(fam.wmax) + 1) * fam.disc
    t0 = time.perf_counter()
    fam.precompute_ports(np.column_stack([wc] + [wc] * 8))
    print(f"bench: {len(wc)} port modes in {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    os.makedirs(DESIGN_DATA, exist_ok=True)
    if args.design_iters > 0:
        t0 = time.perf_counter()
        net = design.fit(fam, train_p, train_t, steps=args.fit_steps)
        t_fit = time.perf_counter() - t0
        loop = design.SurrogateLoop(fam, net, train_p, train_t, spec=spec, seed=WL.TAPER_SEED + 1, min_hit=0.0)
        print("bench: surrogate loop", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for it in range(args.design_iters):
            loop.step()
            print(f"bench: iteration {it + 1}: {len(loop.accepted)} accepted of {loop.verified} verified, "
                  f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        t_sur = time.perf_counter() - t0
        err = loop.surrogate_error()
        net.free()
        acc, strict = len(loop.accepted), int((err <= 1e-3).sum())
        np.savez(os.path.join(DESIGN_DATA, "cfg5_accepted.npz"), p=np.array(loop.accepted), spec=spec,
                 fdtd=np.array(loop.spec_fdtd), sur=np.array(loop.spec_sur))
        part = {"accepted": acc, "accepted_strict": strict, "verified": loop.verified, "iterations": loop.iterations,
                "seconds": round(t_sur, 1), "designs_per_hour": round(acc / t_sur * 3600, 1),
                "designs_per_hour_strict": round(strict / t_sur * 3600, 1), "training_set_fdtd_s": round(t_data, 1),
                "training_s": round(t_fit, 1),
                "designs_per_hour_including_training_set": round(acc / (t_sur + t_data + t_fit) * 3600, 1),
                "surrogate_max_dT_on_accepted": {"median": round(float(np.median(err)), 4) if acc else None,
                                                 "max": round(float(err.max()), 4) if acc else None,
                                                 "within_1e-3": strict},
                "train_designs": len(train_p), "fit_steps": args.fit_steps, "spec": spec,
                "screened_per_iteration": loop.pool, "verified_per_iteration": loop.verify}
        with open(os.path.join(DESIGN_DATA, "cfg5_surrogate.json"), "w") as fh:
            json.dump(part, fh)
    if args.plain_seconds > 0:
        plain_path = os.path.join(DESIGN_DATA, "cfg5_plain.json")
        prev = json.load(open(plain_path)) if args.plain_append and os.path.exists(plain_path) else None
        # --plain-append continues a saved part with the starts after its own (the seed advances)
        rng = np.random.default_rng(WL.TAPER_SEED + 7 + (prev["starts"] if prev else 0))
        print("bench: plain loop", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        plain_ok, iters, starts = 0, 0, 0
        plain_acc = []
        while time.perf_counter() - t0 < args.plain_seconds:
            p0 = fam.random(rng, 1)[0]
            p, hist = design.adjoint_optimize(fam, p0, 12, spec)
            starts += 1
            iters += len(hist)
            if hist[-1] >= spec:
                plain_ok += 1
                plain_acc.append(p)
            print(f"bench: plain start {starts}: T {hist[0]:.3f} -> {hist[-1]:.3f} in {len(hist)} iterations, "
                  f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
        t_plain = time.perf_counter() - t0
        part = design.plain_part(plain_ok, starts, iters, t_plain, spec, plain_acc)
        if prev:
            part = design.merge_plain_parts(prev, part)
        with open(plain_path, "w") as fh:
            json.dump(part, fh)
    paths = [os.path.join(DESIGN_DATA, f"cfg5_{w}.json") for w in ("surrogate", "plain")]
    if not all(os.path.exists(q) for q in paths):
        print("bench: one loop's part saved; the line is printed once both parts exist", file=sys.stderr, flush=True)
        return
    sur, plain = (json.load(open(q)) for q in paths)
    rate, rate_strict, rate_plain = sur["designs_per_hour"], sur["designs_per_hour_strict"], plain["designs_per_hour"]
    out = {"metric": METRIC, "value": rate, "unit": "verified taper designs/hour", "n_gpus": 1, "steps": fam.steps,
           "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 fields, f64 monitors, bf16 surrogate", "data": "synthetic",
           "config": {"workload": "taper family of the batched 64-taper set (20 nm cells, 200x100 cross-section, 8-cell CPML, lengths U[3,15] "
                                  "um, 8 spline widths U[0.4,3.2] um, 6000 steps, 16 wavelengths 1.50-1.60 um, TE0 "
                                  "ports from the mode solver); surrogate MLP 9-256x4-32 GELU (BF16 MFMA, FP32 "
                                  "accumulate) trained on the FDTD training set; 10^5 screened and 64 verified by "
                                  "batched FDTD per iteration",
                      "train_designs": sur["train_designs"], "fit_steps": sur["fit_steps"],
                      "spec": f"band |t|^2 >= {sur['spec']:.4f} (95th percentile of the training set)",
                      "screened_per_iteration": sur["screened_per_iteration"],
                      "verified_per_iteration": sur["verified_per_iteration"]},
           "design_loop": {"surrogate": sur, "plain": {k: v for k, v in plain.items() if k != "accepted_designs"},
                           "ratio": round(rate / rate_plain, 2) if rate_plain else None,
                           "ratio_strict": round(rate_strict / rate_plain, 2) if rate_plain else None}}
    print(json.dumps(out), flush=True)


def design_oracle_check(args, n=3):
    """FP64 oracle forward runs (CPU) of the first n accepted designs of the last design campaign
    (DESIGN_DATA/cfg5_accepted.npz): the GPU FDTD transmission of each accepted design against FP64."""
    from oracle import taper_oracle

    design = importlib.import_module("3d_fdtd_amd.design")
    fam = design_family(design)
    z = np.load(os.path.join(DESIGN_DATA, "cfg5_accepted.npz"))
    p = z["p"][:n]
    out = []
    for q, pp in enumerate(p):
        t0 = time.perf_counter()
        t_cpu = taper_oracle.transmission(**fam.run_inputs(pp))
        T_cpu = np.abs(t_cpu) ** 2
        out.append({"design": [round(float(v), 4) for v in pp], "oracle_s": round(time.perf_counter() - t0, 1),
                    "max_abs_dT_vs_gpu": float(np.abs(T_cpu - z["fdtd"][q]).max()),
                    "band_T_fp64": float(T_cpu.mean())})
        print(json.dumps(out[-1]), flush=True)
    return out


def bench_gradient(fd, args):
    """The 3-D taper with its DFT-domain permittivity gradient: a linear Si taper 0.5 -> 3.0 um
    over 10 um between strip ports (8 spline control widths), 620 x 200 x 100 cells of 20 nm
    (library axes 200 x 100 x 620), 10-cell CPML, TE0 of each port from the mode solver.  One
    forward run (6000 steps, FP64 DFT box over the taper's core rows at 1.55 um) and one adjoint
    run of the output monitor (6000 steps) give d|t|^2/d eps over the box and, through the edge
    cells, d|t|^2/d w_i; a central difference of two more forward runs checks one width.
    value = cell-updates/s over the two gradient runs (device time); roofline = the split kernels
    in a separate 200-step window with every launch timed."""
    design = importlib.import_module("3d_fdtd_amd.design")
    steps = args.steps if args.steps != N_RUN else 6000
    fam = design.TaperFamily(disc=DISC, I=200, J=100, core_cells=11, npml=10, port=48, steps=steps,
                             lam=(1.55, 1.55), nfreq=1)
    p = np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]])
    ep, K = fam.permittivity(p)
    I, J = fam.I, fam.J
    dt, w0 = fam.dt, 2 * math.pi / 1.55
    ci, cj, ks, km = I // 2, J // 2, fam.npml + 1, K - fam.npml - 2
    print("bench: solving the port modes", file=sys.stderr, flush=True)
    m_in, mi_in, mj_in, n_in = WL.port_mode(fam, p, ks, w0)
    m_out, mi_out, mj_out, n_out = WL.port_mode(fam, p, km, w0)
    j0 = J // 2 - fam.core // 2
    klo, khi = fam.npml + fam.port + 2, K - fam.npml - fam.port - 2
    box = (0, I, j0, j0 + fam.core, klo, khi)
    nb = I * fam.core * (khi - klo)
    om = np.array([w0])

    # grids rasterised on the device from the taper's extruded table (the same permittivity as
    # the host raster ep, which the CPU baseline uses)
    def forward(pp, with_box):
        ex, _ = fam.extruded(pp)
        g = fd.DevGrid(I, J, K, fam.disc, None, extruded=ex)
        g.cpml((fam.npml,) * 6, 3, 1e-8, 1.0, 0.05)
        g.add_source(ci, cj, ks, m_in, mi_in, mj_in, fam.disc, fam.amp, fam.phase)
        g.mon_in = g.add_overlap(ci, cj, ks + 3, m_in, mi_in, mj_in, fam.disc, om)
        g.mon_out = g.add_overlap(ci, cj, km, m_out, mi_out, mj_out, fam.disc, om)
        g.box = g.add_dft_box(box, 7, om) if with_box else None
        return g

    def transmission(g):
        a_in, _ = g.overlap(g.mon_in, 1)
        a_out, _ = g.overlap(g.mon_out, 1)
        return a_out[0] / a_in[0], a_in[0]

    print(f"bench: forward run ({steps} steps, {I * J * K / 1e6:.1f} M cells)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g = forward(p, True)  # device raster, material ids and tables, CPML, source and monitors
    fd.dev_check(fd.lib().fdtd_dev_synchronize(g.d), "synchronize")
    t_setup = time.perf_counter() - t0
    g.timer_start()
    g.run(dt, 0, steps)
    t_fwd = g.timer_stop() / 1e3
    t, a_in = transmission(g)
    print("bench: adjoint run", file=sys.stderr, flush=True)
    ga = fd.DevGrid(I, J, K, fam.disc, None, extruded=fam.extruded(p)[0])
    ga.cpml((fam.npml,) * 6, 3, 1e-8, 1.0, 0.05)
    ga.add_adjoint_source(ci, cj, km, m_out, mi_out, mj_out, fam.disc, w0 * dt, fam.amp, fam.phase)
    aid = ga.add_dft_box(box, 7, om)
    ga.timer_start()
    ga.run(dt, 0, steps, 1.0)
    t_adj = ga.timer_stop() / 1e3
    t0 = time.perf_counter()
    weight = 2.0 * np.conj(t) * design.adjoint_scale(fam.amp, w0, w0, dt) / a_in
    grad_eps = g.gradient(g.box, ga, aid, 0, weight, nb)
    t_grad = time.perf_counter() - t0
    g.free()
    ga.free()
    dT_dw = fam.width_gradient(p, K, grad_eps.reshape(khi - klo, fam.core, I), klo)
    print("bench: finite-difference check of w_4", file=sys.stderr, flush=True)
    d = 0.01
    T_pm = []
    for sgn in (1, -1):
        pp = p.copy()
        pp[3] += sgn * d
        gf = forward(pp, False)
        gf.run(dt, 0, steps)
        T_pm.append(abs(transmission(gf)[0]) ** 2)
        gf.free()
    fd_w4 = (T_pm[0] - T_pm[1]) / (2 * d)
    print("bench: kernel-timing window", file=sys.stderr, flush=True)
    gk = fd.DevGrid(I, J, K, fam.disc, None, extruded=fam.extruded(p)[0])
    gk.cpml((fam.npml,) * 6, 3, 1e-8, 1.0, 0.05)
    gk.add_source(ci, cj, ks, m_in, mi_in, mj_in, fam.disc, fam.amp, fam.phase)
    gk.run(dt, 0, 800)
    gk.kernel_timing(1)
    gk.run(dt, 800, 200)
    (e_us, h_us), _ = gk.kernel_times()
    gk.kernel_timing(0)
    gk.free()
    n = I * J * K
    dom = max(e_us, h_us)
    achieved = HALF_STEP_BYTES * n / (dom * 1e-6) / 1e9
    dom_name = "yee_update_E" if e_us >= h_us else "yee_update_H"
    traffic = read_traffic("yee_traffic_cfg3_taper.json")
    out = {"metric": METRIC, "value": 2 * steps * n / (t_fwd + t_adj), "unit": "cell-updates/s", "n_gpus": 1,
           "steps": 2 * steps, "warmup": 0, "ms_per_step": (t_fwd + t_adj) * 1e3 / (2 * steps),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "linear Si taper 0.5 -> 3.0 um over 10 um between strip ports, 8 spline control "
                                  "widths, 620x200x100 cells of 20 nm, 10-cell CPML, TE0 port modes from the mode "
                                  "solver; forward + adjoint runs with an FP64 DFT box over the taper's core rows at "
                                  "1.55 um and the DFT-domain eps-gradient",
                      "grid_lib_axes": [I, J, K], "steps_per_run": steps, "dft_box_cells": nb},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic.get(dom_name) if traffic else None,
                        "kernel": dom_name,
                        "kernel_avg_us": round(dom, 2), "e_us": round(e_us, 2), "h_us": round(h_us, 2),
                        "algorithmic_bytes_per_launch": HALF_STEP_BYTES * n,
                        "step_frac_48B": round(STEP_BYTES * n / ((e_us + h_us) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
           "gradient": {"setup_s": round(t_setup, 3), "forward_s": round(t_fwd, 3), "adjoint_s": round(t_adj, 3), "gradient_s": round(t_grad, 4),
                        "T": abs(t) ** 2, "n_eff_in": round(float(n_in.real), 4), "n_eff_out": round(float(n_out.real), 4),
                        "dT_dw_per_um": [round(float(v), 6) for v in dT_dw],
                        "fd_check_w4": {"adjoint": round(float(dT_dw[3]), 6), "central_difference": round(float(fd_w4), 6),
                                        "delta_um": d}}}
    if not args.no_cpu:
        out["cpu_baseline"] = gradient_cpu_baseline(fam, ep, K)
    print(json.dumps(out), flush=True)


def gradient_cpu_baseline(fam, ep, K, target_s=10.0):
    """FP64 oracle CPML steps on the same grid (host threads), about target_s of them."""
    from oracle import oracle_py as O

    threads = int(os.environ.get("FDTD3D_CPU_THREADS", physical_cores()[0]))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    g = O.Grid(fam.I, fam.J, K, fam.disc)
    g.ep[:] = ep
    cp = O.Cpml(g, (fam.npml,) * 6, fam.dt, order=3, R0=1e-8, kappa_max=1.0, alpha_max=0.05)
    g.f[:] = np.random.default_rng(0).uniform(-1, 1, g.f.size)
    t0 = time.perf_counter()
    O.step_cpml(g, cp, fam.dt)
    steps = max(2, min(1000, int(target_s / max(time.perf_counter() - t0, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(steps):
        O.step_cpml(g, cp, fam.dt)
    el = time.perf_counter() - t0
    return {"value": fam.I * fam.J * K * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} FP64 oracle CPML steps of the {fam.I}x{fam.J}x{K} grid, {el:.1f} s, "
                      f"OMP_NUM_THREADS={threads}"}


def bench_shard(fd, args):
    """One GPU's share of the 620x200x100 taper when the grid is split into 8 z slabs (library
    axes 200 x 100 x 620): a middle shard of 78 owned planes plus its two ghost planes (200 x 100 x
    80), CPML on the lateral faces only, paired with its upper neighbour shard on the same GPU for
    the halo planes (fdtd_dev_copy_planes: Ex, Ey down after the E update, Hx, Hy up after the H
    update, 2 x 2 x 80 KB per step).  value = the shard's owned cell-updates/s from its own E and H
    kernel durations (every launch event-timed); the halo copies are timed separately.  Across
    GPUs the planes would travel over xGMI (RCCL send/recv or peer copies); the device-to-device
    copy stands in for them here."""
    design = importlib.import_module("3d_fdtd_amd.design")
    fam = design.TaperFamily(disc=DISC, I=200, J=100, core_cells=11, npml=10, port=48, steps=10,
                             lam=(1.55, 1.55), nfreq=1)
    ep, K = fam.permittivity(np.concatenate([np.linspace(0.5, 3.0, 8), [10.0]]))
    I, J = fam.I, fam.J
    ep = ep.reshape(K, J, I, 3)
    owned = 78
    k0 = 3 * owned + 1  # shard 3 of 8: global planes [k0, k0 + 78), ghosts k0 - 1 and k0 + 78
    dt = fam.dt
    rng = np.random.default_rng(5)
    npml = (fam.npml, fam.npml, fam.npml, fam.npml, 0, 0)
    shards = []
    for lo in (k0 - 1, k0 + owned - 1):
        g = fd.DevGrid(I, J, owned + 2, DISC, ep[lo:lo + owned + 2].reshape(-1))
        g.cpml(npml, 3, 1e-8, 1.0, 0.05)
        g.set_fields(rng.uniform(-1, 1, 6 * I * J * (owned + 2)))
        shards.append(g)
    a, b = shards
    L = fd.lib()
    EXY, HXY = 0b000011, 0b011000

    def step():
        for g in shards:
            fd.dev_check(L.fdtd_dev_update_E_cpml(g.d, dt), "E")
        fd.dev_check(L.fdtd_dev_copy_planes(a.d, owned + 1, b.d, 1, EXY), "halo E")
        for g in shards:
            fd.dev_check(L.fdtd_dev_update_H_cpml(g.d, dt), "H")
        fd.dev_check(L.fdtd_dev_copy_planes(b.d, 0, a.d, owned, HXY), "halo H")

    S = args.steps if args.steps != N_RUN else 400
    for _ in range(50):
        step()
    for g in shards:
        fd.dev_check(L.fdtd_dev_synchronize(g.d), "sync")
    t0 = time.perf_counter()
    for _ in range(S):
        step()
    for g in shards:
        fd.dev_check(L.fdtd_dev_synchronize(g.d), "sync")
    pair_s = (time.perf_counter() - t0) / S
    # kernel durations of the shard's own launches, without the copies between them (a launch
    # queued behind a cross-stream wait takes its start stamp early)
    a.kernel_timing(1)
    for _ in range(min(S, 200)):
        fd.dev_check(L.fdtd_dev_update_E_cpml(a.d, dt), "E")
        fd.dev_check(L.fdtd_dev_update_H_cpml(a.d, dt), "H")
    (e_us, h_us), cnt = a.kernel_times()
    a.kernel_timing(0)
    t0 = time.perf_counter()
    for _ in range(S):
        fd.dev_check(L.fdtd_dev_copy_planes(a.d, owned + 1, b.d, 1, EXY), "halo E")
        fd.dev_check(L.fdtd_dev_copy_planes(b.d, 0, a.d, owned, HXY), "halo H")
    for g in shards:
        fd.dev_check(L.fdtd_dev_synchronize(g.d), "sync")
    copy_us = (time.perf_counter() - t0) / S * 1e6
    for g in shards:
        g.free()
    cells = I * J * owned
    dom = max(e_us, h_us)
    achieved = HALF_STEP_BYTES * cells / (dom * 1e-6) / 1e9
    out = {"metric": METRIC, "value": cells / ((e_us + h_us) * 1e-6), "unit": "cell-updates/s", "n_gpus": 1,
           "steps": S, "warmup": 50, "ms_per_step": (e_us + h_us) / 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "one of 8 z-slab shards of the 620x200x100 taper: 78 owned planes + 2 ghost "
                                  "planes (library 200x100x80), lateral 10-cell CPML, halo planes exchanged with "
                                  "a neighbour shard on the same GPU",
                      "owned_cells": cells, "halo_bytes_per_step": 2 * 2 * I * J * 4},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "yee_update_E" if e_us >= h_us else "yee_update_H",
                        "kernel_avg_us": round(dom, 2), "e_us": round(e_us, 2), "h_us": round(h_us, 2),
                        "timed_launches": int(min(cnt)), "algorithmic_bytes_per_launch": HALF_STEP_BYTES * cells},
           "shard": {"pair_step_us_two_shards_one_gpu": round(pair_s * 1e6, 1),
                     "halo_copy_pair_us_device_to_device": round(copy_us, 1),
                     "full_grid_cells_per_shard_fraction": round(owned / K, 4)}}
    print(json.dumps(out), flush=True)


def dropin_setup(L, TF, complex_t, steps):
    """The reference driver's forward-run set-up (3d_fdtd.c:30-92) on library L: 100 x 100 x 600
    cells of 20 nm, SiO2 background with nitride / BPSG / SiCOH top layers, input taper and
    waveguide, a 25-segment adjoint taper inserted at planes 150..400, output waveguide and taper.
    Its mode files are not part of the reference, so Gaussian profiles (60 x 40 cells) of the same
    role stand in for the launch, output and unwanted modes."""
    disc, omega = 0.02, 2 * math.pi / 1.310
    tf = L.timefield3d_new(100, 100, 600, disc)
    L.timefield3d_set_background(tf, 1.447 ** 2, 1.0)
    for ep, bottom in ((1.93 ** 2, 55), (1.47 ** 2, 59), (1.39 ** 2, 64)):
        L.timefield3d_set_top_half_plane(tf, ep, 1.0, bottom)
    L.timefield3d_set_taper(tf, 3.51 ** 2, 1.0, 50, 50, 41, 59, 0, 95, 46, 54)
    L.timefield3d_set_waveguide(tf, 3.51 ** 2, 41, 59, 95, 150, 46, 54)
    seg = 25
    left = (C.c_int * (seg + 1))(*[41 - i * 9 // seg for i in range(seg + 1)])
    right = (C.c_int * (seg + 1))(*[59 + i * 9 // seg for i in range(seg + 1)])
    taper = L.adjoint3d_taper_new(3.51 ** 2, 1.0, seg, 150, 400, left, right, 46, 54, steps)
    L.adjoint3d_taper_insert(tf, taper)
    L.timefield3d_set_waveguide(tf, 3.51 ** 2, 32, 68, 400, 455, 46, 54)
    L.timefield3d_set_taper(tf, 3.51 ** 2, 1.0, 32, 68, 50, 50, 455, 600, 46, 54)
    modes = [L.mode_gaussian(60, 40, disc, complex_t(1.447 ** 2, 0), complex_t(1.0, 0), omega, w, 0)
             for w in (6.0, 12.0, 9.0)]
    return tf, taper, modes, omega, 0.99 * disc / math.sqrt(3)


def dropin_steps(L, tf, taper, modes, omega, dt, complex_t, n0, n):
    """n steps of the reference driver's forward loop (3d_fdtd.c:104-127), call for call."""
    launch, output, unwanted = modes
    acc = [complex_t(0, 0) for _ in range(4)]
    for i in range(n0, n0 + n):
        amp = math.exp(-(i - 300) ** 2 / (2 * 150.0 ** 2)) if i < 1000 else 0.0
        phase = i * dt * omega
        L.timefield3d_mode_source(tf, 50, 50, 125, launch, phase, amp, 0)
        L.timefield3d_apply_timestep_omp_full(tf, dt)
        L.timefield3d_update_fourier_fields_y_slice_omp(tf, phase, 50)
        L.timefield3d_mode_overlap(tf, 50, 50, 130, launch, phase, 0, C.byref(acc[0]))
        L.timefield3d_mode_overlap(tf, 50, 50, 130, launch, phase, 1, C.byref(acc[1]))
        L.timefield3d_mode_overlap(tf, 50, 50, 425, output, phase, 0, C.byref(acc[2]))
        L.timefield3d_mode_overlap(tf, 50, 50, 425, unwanted, phase, 0, C.byref(acc[3]))
        L.adjoint3d_taper_store_forward_fields(taper, tf, i)
    return acc


def bench_dropin(fd, args):
    """The reference driver's forward loop through the drop-in C surface, call for call: per step
    a mode source, the full step (Mur faces), the y-slice DFT, four mode overlaps (each returns a
    value, so the step synchronises four times) and the taper's forward-history store, on the
    driver's 100 x 100 x 600 grid.  value = cell-updates/s of that loop (wall clock; the
    per-step returns make host time part of it).  cpu_baseline: the same calls on the compiled
    reference (OpenMP, host threads)."""
    L = fd.lib()
    W, S = 20, args.steps if args.steps != N_RUN else 300
    tf, taper, modes, omega, dt = dropin_setup(L, fd.TimeField, fd.Complex, W + S)
    dropin_steps(L, tf, taper, modes, omega, dt, fd.Complex, 0, W)
    t0 = time.perf_counter()
    acc = dropin_steps(L, tf, taper, modes, omega, dt, fd.Complex, W, S)
    el = time.perf_counter() - t0
    cells = 100 * 100 * 600
    print(f"bench: drop-in loop {el / S * 1e3:.3f} ms/step, |launched| {abs(complex(acc[0].re, acc[0].im)):.3e}",
          file=sys.stderr, flush=True)
    out = {"metric": METRIC, "value": cells * S / el, "unit": "cell-updates/s", "n_gpus": 1, "steps": S, "warmup": W,
           "ms_per_step": el / S * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 fields, f64 monitors", "data": "synthetic",
           "config": {"workload": "the reference driver's forward loop (3d_fdtd.c:104-127) through the drop-in C "
                                  "surface: 100x100x600 cells of 20 nm, Mur faces, mode source, y-slice DFT, four "
                                  "mode overlaps and the taper history store per step; Gaussian stand-ins for the "
                                  "driver's mode files"}}
    if not args.no_cpu:
        from oracle import ref_py

        if ref_py.available():
            threads = int(os.environ.get("FDTD3D_CPU_THREADS", physical_cores()[0]))
            os.environ["OMP_NUM_THREADS"] = str(threads)
            R = ref_py.lib()
            rtf, rtaper, rmodes, omega, dt = dropin_setup(R, ref_py.TimeField, ref_py.Complex, 2000)
            dropin_steps(R, rtf, rtaper, rmodes, omega, dt, ref_py.Complex, 0, 2)
            t0 = time.perf_counter()
            n = 0
            while time.perf_counter() - t0 < 12.0 and n < 2000 - 2:
                dropin_steps(R, rtf, rtaper, rmodes, omega, dt, ref_py.Complex, 2 + n, 1)
                n += 1
            cel = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": cells * n / cel, "unit": "cell-updates/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": f"{n} steps of the same call sequence on the compiled reference "
                                             f"(FP64, OpenMP), {cel:.1f} s, OMP_NUM_THREADS={threads}"}
    print(json.dumps(out), flush=True)


def bench_surrogate(fd, args):
    """The transmission surrogate's hand-written BF16 matrix-core kernels: full-batch Adam steps on
    4096 samples and screening of 10^5 candidates (the design loop's shapes: 9 -> 256 x 4 -> 32,
    GELU).  value = candidates screened per second (device time on the network's stream, host
    transfers of inputs and outputs included); roofline = the GEMM launches' arithmetic over their
    own dispatch-recorded durations, against the dense BF16 matrix-core peak."""
    design = importlib.import_module("3d_fdtd_amd.design")
    dims = (9, 256, 4, 32)
    rng = np.random.default_rng(WL.TAPER_SEED)
    n_train, n_pool = 4096, 100000
    x = rng.uniform(-1, 1, (n_train, dims[0])).astype(np.float32)
    t = np.sin(x @ rng.normal(0, 0.6, (dims[0], dims[3]))).astype(np.float32)
    pool = rng.uniform(-1, 1, (n_pool, dims[0])).astype(np.float32)
    net = fd.Mlp(*dims, params=design.he_uniform(dims, 0))
    net.train(x, t, 20)
    steps = args.steps if args.steps != N_RUN else 200
    net.timer_start()
    loss0, loss1 = net.train(x, t, steps)
    train_ms = net.timer_stop() / steps
    net.gemm_timing(True)
    net.train(x, t, 20)
    tl, tms, tfl = net.gemm_times()
    net.forward(pool)
    reps = 20
    net.gemm_timing(True)
    net.timer_start()
    for _ in range(reps):
        net.forward(pool)
    fwd_ms = net.timer_stop() / reps
    fl, fms, ffl = net.gemm_times()
    net.gemm_timing(False)
    net.free()
    tflops = ffl / (fms / 1e3) / 1e12
    # algorithmic bytes of one screening forward: each layer reads its bf16 input rows and weights
    # and writes bf16 activations (FP32 outputs for the last layer); 41.4 GFLOP over ~426 MB is
    # ~97 FLOP/B, below the bf16 ridge (2.5 PFLOP/s / 8 TB/s = 312 FLOP/B): the GEMMs are HBM-bound
    kin = (dims[0] + 7) // 8 * 8
    fins = [kin] + [dims[1]] * dims[2]
    fouts = [dims[1]] * dims[2] + [dims[3]]
    fwd_bytes = sum(n_pool * fi * 2 + fi * fo * 2 + n_pool * fo * (4 if q == len(fins) - 1 else 2)
                    for q, (fi, fo) in enumerate(zip(fins, fouts)))
    achieved = fwd_bytes / (fms / reps / 1e3) / 1e9
    out = {"metric": METRIC, "value": n_pool / (fwd_ms / 1e3), "unit": "screened candidates/s", "n_gpus": 1,
           "steps": steps, "warmup": 20, "ms_per_step": train_ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 operands, fp32 accumulate", "data": "synthetic",
           "config": {"workload": "transmission surrogate MLP 9-256x4-32 GELU: Adam on 4096 samples, screening "
                                  "10^5 candidates", "train_samples": n_train, "screened": n_pool},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "k_gemm_nt (screening forward, 5 launches)",
                        "algorithmic_bytes_per_forward": fwd_bytes, "gemm_ms_per_forward": round(fms / reps, 4),
                        "mfma_tflops": round(tflops, 2), "mfma_frac_of_2500": round(tflops / 2500.0, 5),
                        "train_gemm_tflops": round(tfl / (tms / 1e3) / 1e12, 2)},
           "surrogate": {"adam_step_ms": round(train_ms, 4), "screen_1e5_ms": round(fwd_ms, 4),
                         "loss_before": loss0, "loss_after": loss1}}
    print(json.dumps(out), flush=True)


def cpu_baseline(target_s=12.0):
    """Reference FP64 step on the same grid size, host cores, bounded sample."""
    threads = int(os.environ.get("FDTD3D_CPU_THREADS", physical_cores()[0]))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    dt = 0.99 * DISC / math.sqrt(3)
    from oracle import ref_py

    if ref_py.available():
        L = ref_py.lib()
        tf = L.timefield3d_new(I, J, K, DISC)
        L.timefield3d_set_background(tf, 1.44 ** 2, 1.0)
        L.timefield3d_set_prism_permittivity(tf, 3.48 ** 2, I // 2 - 12, I // 2 + 13, J // 2 - 5, J // 2 + 6, 0, K)
        A = ref_py.tf_arrays(tf)
        A["fields"][:] = np.random.default_rng(0).uniform(-1, 1, A["fields"].size)

        def step():
            L.timefield3d_apply_timestep_omp_full(tf, dt)

        kind = "reference"
        cleanup = lambda: L.timefield3d_free(tf)  # noqa: E731
    else:
        from oracle import oracle_py as O

        o = O.Grid(I, J, K, DISC)
        o.f[:] = np.random.default_rng(0).uniform(-1, 1, o.f.size)

        def step():
            o.step(dt)

        kind = "port"
        cleanup = lambda: None  # noqa: E731
    step()  # warm
    t0 = time.perf_counter()
    step()
    est = time.perf_counter() - t0
    n = max(2, min(2000, int(target_s / max(est, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(n):
        step()
    el = time.perf_counter() - t0
    cleanup()
    rate = I * J * K * n / el
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": f"{n} steps of timefield3d_apply_timestep_omp_full (FP64, Mur) on the {I}x{J}x{K} grid, "
                      f"{el:.1f} s, OMP_NUM_THREADS={threads}"}


def calibrate_copy(mib, reps=8):
    """A plain read+write elementwise kernel of known size (torch.neg over `mib` MiB of FP32):
    what rocprofv3's FETCH_SIZE / WRITE_SIZE read for it scales those counters to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "kernel": "torch.neg (elementwise FP32)", "launches": reps,
                      "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4}), flush=True)


def read_traffic(name="yee_traffic_cfg2.json"):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh)


def bench_waveguide(fd, args):
    """The strip waveguide (one grid, device-resident run API): cell-updates/s of K steps."""
    L = fd.lib()
    W, S = args.warmup, args.steps
    g, dt, md, mons, omegas = WL.strip_waveguide(W + S + min(400, max(S // 4, 20)))
    if args.rows is not None or args.kchunk is not None:
        rows = 0 if args.rows is None else args.rows
        kchunk = args.kchunk if args.kchunk is not None else (0 if rows == 0 else 8)
        fd.check(L.fdtd3d_ext_set_step_kernel(g.tf, rows, kchunk), "step kernel")
    print(f"bench: warmup {W} steps", file=sys.stderr, flush=True)
    fd.check(L.fdtd3d_ext_run(g.tf, dt, 0, W, 0.0), "warmup run")
    fd.check(L.fdtd3d_ext_synchronize(g.tf), "sync")
    print(f"bench: timing {S} steps", file=sys.stderr, flush=True)
    fd.check(L.fdtd3d_ext_timer_start(g.tf), "timer")
    t0 = time.perf_counter()
    fd.check(L.fdtd3d_ext_run(g.tf, dt, W, S, 0.0), "timed run")
    ms = C.c_double()
    fd.check(L.fdtd3d_ext_timer_stop(g.tf, C.byref(ms)), "timer")
    wall = time.perf_counter() - t0
    # Kernel durations: a second timed window of the same run in which every E / H launch carries
    # a start/stop event pair recorded by its own dispatch.  Sampling only some launches of a
    # back-to-back stream reads 1.5-2 us high (a queued dispatch takes its start timestamp before
    # the previous kernel has drained); with every launch timed the launches are spaced by the
    # host's event work, and the durations agree with rocprofv3's to about 1 %.  The throughput
    # window above carries no events.
    KW = min(400, max(S // 4, 20))
    fd.check(L.fdtd3d_ext_kernel_timing(g.tf, 1), "kernel timing")
    fd.check(L.fdtd3d_ext_run(g.tf, dt, W + S, KW, 0.0), "kernel-timing run")
    counts, sums = (C.c_int * 3)(), (C.c_double * 3)()
    fd.check(L.fdtd3d_ext_kernel_times(g.tf, counts, sums), "kernel times")
    fd.check(L.fdtd3d_ext_kernel_timing(g.tf, 0), "kernel timing off")
    elapsed_s = max(ms.value / 1e3, wall)
    cells = I * J * K
    value = cells * S / elapsed_s
    # transmission |t|^2 at 1.55 um as a sanity print
    cum = (fd.Complex * len(omegas))()
    launched = (fd.Complex * len(omegas))()
    L.fdtd3d_ext_run_overlap(g.tf, mons[2], cum, None)
    L.fdtd3d_ext_run_overlap(g.tf, mons[0], launched, None)
    tr = [abs(complex(cum[q].re, cum[q].im) / complex(launched[q].re, launched[q].im)) ** 2
          if (launched[q].re or launched[q].im) else float("nan") for q in range(len(omegas))]
    print(f"bench: |t|^2 over 1.50-1.60 um: {min(tr):.4f} .. {max(tr):.4f}", file=sys.stderr, flush=True)
    avg = [sums[w] / counts[w] if counts[w] else 0.0 for w in range(3)]
    if counts[2]:  # fused step kernel: a whole cell-update per cell and launch
        dom_name, dom_ms, dom_bytes = "yee_step_fused", avg[2], STEP_BYTES
        step_ms = avg[2]
    else:  # split pair: each launch advances half a cell-update
        dom_name, dom_ms = ("yee_update_E", avg[0]) if avg[0] >= avg[1] else ("yee_update_H", avg[1])
        dom_bytes = HALF_STEP_BYTES
        step_ms = avg[0] + avg[1]
    achieved = dom_bytes * cells / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    traffic = read_traffic("yee_traffic_cfg2.json")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic.get(dom_name) if traffic else None,
            "kernel": dom_name, "kernel_avg_us": round(dom_ms * 1e3, 2),
            "algorithmic_bytes_per_launch": dom_bytes * cells,
            "timed_launches": int(max(counts[0], counts[1], counts[2])),
            "kernel_timing_window_steps": KW,
            "step_kernels_avg_us": round(step_ms * 1e3, 2),
            "frac_of_value_48B": round(STEP_BYTES * value / 1e9 / HBM_PEAK_GBS, 4)}
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": S, "warmup": W,
           "ms_per_step": elapsed_s * 1e3 / S, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "Si strip waveguide 240x120x100 (x-propagation, 20 nm cells), 10-cell CPML, "
                                  "TE0 mode source from the mode solver, FP32 fields, 3 FP64 overlap monitors x 16 "
                                  "wavelengths",
                      "grid_lib_axes": [I, J, K], "steps_per_run": N_RUN},
           "roofline": roof}
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    L.mode_free(md)
    g.free()
    print(json.dumps(out), flush=True)



def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 20 for the batched tapers, 4000 for the waveguide)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps before the timed ones")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--rows", type=int, default=None, help="fused step rows per lane (0 = split kernels)")
    ap.add_argument("--kchunk", type=int, default=None,
                    help="z planes per wavefront (fused step) or per thread (split kernels, 0 = automatic)")
    ap.add_argument("--workload", choices=("tapers", "waveguide", "design", "surrogate", "gradient", "shard",
                                           "dropin"),
                    default="tapers",
                    help="tapers: the batched 64-taper set (default, the largest configuration); waveguide: the "
                         "strip waveguide; design: verified taper designs/hour, surrogate loop vs plain adjoint "
                         "loop; surrogate: the MLP's matrix-core training and screening kernels; gradient: the "
                         "620x200x100 taper's forward + adjoint runs and DFT eps-gradient; shard: one GPU's z-slab "
                         "share of that taper with its halo planes; dropin: the reference driver's per-step call "
                         "sequence through the drop-in C surface")
    ap.add_argument("--train", type=int, default=4096, help="design workload: FDTD training designs")
    ap.add_argument("--design-phase", choices=("loop", "data", "oracle"), default="loop",
                    help="design workload: generate a training-set chunk, run the loops, or check accepted "
                         "designs against the FP64 oracle (CPU)")
    ap.add_argument("--chunk", type=int, default=0, help="design data: chunk index")
    ap.add_argument("--chunks", type=int, default=5, help="design data: chunks of the training set")
    ap.add_argument("--design-iters", type=int, default=10, help="design workload: surrogate-loop iterations")
    ap.add_argument("--plain-seconds", type=float, default=360.0, help="design workload: plain-loop wall time")
    ap.add_argument("--plain-append", action="store_true",
                    help="design workload: add this call's plain-loop starts to the saved plain part")
    ap.add_argument("--fit-steps", type=int, default=1500, help="design workload: surrogate training steps")
    ap.add_argument("--tapers", type=int, default=64, help="number of taper grids of the batched workload")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel over MIB MiB (HBM counter calibration)")
    # child-process CPU leg of the batched workload's baseline (no GPU use)
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-taper", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-sample", type=float, default=3.0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg(args)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "geometry for batched runs or by z slabs with halo planes for one large grid, "
              "fdtd_dev_copy_planes; no multi-GPU run exists)", file=sys.stderr)
        sys.exit(2)
    fd = importlib.import_module("3d_fdtd_amd")
    fd.lib()
    if args.workload == "tapers":
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 5 if args.warmup is None else args.warmup
        bench_tapers(fd, args)
        return
    args.warmup = 100 if args.warmup is None else args.warmup
    if args.steps is None:
        args.steps = N_RUN
    if args.workload == "design":
        if args.design_phase == "data":
            design_data(fd, args)
        elif args.design_phase == "oracle":
            design_oracle_check(args)
        else:
            bench_design(fd, args)
    elif args.workload == "surrogate":
        bench_surrogate(fd, args)
    elif args.workload == "gradient":
        bench_gradient(fd, args)
    elif args.workload == "shard":
        bench_shard(fd, args)
    elif args.workload == "dropin":
        bench_dropin(fd, args)
    else:
        bench_waveguide(fd, args)


if __name__ == "__main__":
    main()
