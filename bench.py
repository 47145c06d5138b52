#!/usr/bin/env python3
"""Benchmark: Monte Carlo drainage-basin realizations per second on one MI355X.

Workload: synthetic 4096 x 4096 float32 DEM with 10 m cells (seeded generator
in the library: 100 m base + 0.5 % tilt + 40 Gaussian hills + 0.3 m noise,
streams where D8 accumulation on the filled 5x5-smoothed DEM exceeds 2000
cells), error model N(0, 1.8^2) convolved with the Gaussian filter of
practical range 60 m (and, reported beside it, the exponential filter of
practical range 200 m).  One step is one full realization: random field,
convolution, realistic DEM, stream burning, pit filling, flow routing, flow
tracing and border accumulation, with inputs resident in HBM.

Prints one JSON line.  `value` is realizations/s of the Gaussian workload on
one GPU.  Per-stage device times come from HIP events that the library
records on its stream between the stages of every timed realization (no host
synchronisation inside the timed region).  `roofline` covers the dominant
stage: algorithmic bytes per launch / its average launch time (the pit
fill is two kernels, k_pf_persistent and k_pf_resident).  Unless
--no-traffic, `traffic` is the HBM bytes per launch of those kernels from two
rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE) over a short probe
run (`--probe`), scaled by what the same counters read for an elementwise
kernel that reads and writes 512 MiB; otherwise it is null.
`cpu_baseline` times the CPU oracle (a restatement of the reference's OpenMP
CPU path) on a bounded sample, in a subprocess pinned to one thread.
`batched` runs 64 uneven tiles x 100 realizations as graph replays and as
per-tile calls (identical counters required).
`surrogate` judges the packaged border-probability network (with its
Monte Carlo check) against a 1000-realization Monte Carlo of the same DEM:
per-cell deviation of the border probability and wall time of each.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "realizations/s + Mcells/s at 4096² DEM; surrogate vs MC time @ L1≤0.02"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
NODATA = -9999.0
CALIB_BYTES = 512 << 20  # the calibration kernel reads and writes this much

# Algorithmic bytes per cell per realization of each stage at the reference's
# data types (one pass over stage-boundary I/O; iterations of the fixed-point
# stages are not counted).  realistic_burn is one fused kernel: the unfused
# figure of realistic DEM (12) + burning (12) is counted.
STAGE_BYTES = {
    "random_field": 4, "convolution": 16, "realistic_burn": 24, "pit_fill": 8,
    "flow_routing": 6, "flow_tracing": 10, "border": 12,
}
STAGE_NAMES = ["random_field", "convolution", "realistic_burn", "pit_fill", "flow_routing",
               "flow_tracing", "border"]
# The kernel launches each stage's events bracket (the pit fill's resident
# kernel ends at once when the first kernel converged alone).
STAGE_KERNEL = {
    "random_field": ("k_random_field",), "realistic_burn": ("k_realistic_burn_quant",),
    "pit_fill": ("k_pf_persistent", "k_pf_resident"),
    "flow_routing": ("k_fr_d8", "k_fr_persistent", "k_fr_pick"),
    "flow_tracing": ("k_tr_tiles", "k_tr_jump", "k_tr_exits"), "border": ("k_border",),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--exp-steps", type=int, default=20,
                   help="realizations timed with the exponential filter (0 = skip)")
    p.add_argument("--seed", type=int, default=20240601)
    p.add_argument("--cpu-sample", type=int, default=3,
                   help="realizations timed on the CPU oracle (0 = skip)")
    p.add_argument("--no-stage-times", action="store_true")
    p.add_argument("--batch-tiles", type=int, default=64,
                   help="tiles of the batched workload (0: skip)")
    p.add_argument("--batch-realizations", type=int, default=100)
    p.add_argument("--large-size", type=int, default=16384,
                   help="side of the large-tile configuration (exponential PR 200 m, RMSE "
                        "2.5 m): Monte Carlo vs surrogate + windowed check + refinement; 0 skips")
    p.add_argument("--large-realizations", type=int, default=1000)
    p.add_argument("--large-window", type=int, default=1024)
    p.add_argument("--large-check", type=int, default=8,
                   help="realizations of the large tile's windowed check")
    p.add_argument("--large-refine", type=int, default=200)
    p.add_argument("--large-weights",
                   default=os.path.join(ROOT, "cageo_amd", "weights", "surrogate_exp_b32.bin"),
                   help="surrogate weights for the exponential filter (.bin)")
    p.add_argument("--surrogate-realizations", type=int, default=1000,
                   help="Monte Carlo size the surrogate is judged against (0: skip)")
    p.add_argument("--traffic-out", default=None,
                   help="keep the counter passes' CSV files in this directory")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 counter passes behind roofline.traffic")
    p.add_argument("--probe", action="store_true",
                   help="internal: the short run profiled by the counter passes")
    return p.parse_args()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the all-core figure: the CPUs this process may run on,
    at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(size, sample, seed, threads=1):
    """Time the CPU oracle on `sample` realizations with `threads` OpenMP
    threads in a child."""
    code = f"""
import os, sys, time, json
sys.path.insert(0, {ROOT!r})
import numpy as np, oracle, cageo_amd as ca
dem, st = ca.synth_dem({size}, {size}, {seed}, 10.0, 2000, {NODATA})
r = oracle.filter_radius(60.0, 10.0)
f = ca.create_filter(oracle.GAUSSIAN, 60.0, 10.0, r)
mc = np.zeros(dem.shape, np.float32)
t = time.perf_counter()
for k in range({sample}):
    st_ = oracle.lib().or_realization(dem, st, {size}, {size}, f, r, 1, 0.0, 1.8, {seed}, k,
                                      {NODATA}, {NODATA}, 2, mc)
    assert st_ == 0
dt = time.perf_counter() - t
print(json.dumps({{"seconds": dt, "n": {sample}}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    return res["n"] / res["seconds"], res["seconds"]


def probe(args):
    """Short run for the counter passes: 2 Gaussian realizations at full size
    and one elementwise kernel over 512 MiB of known traffic."""
    import torch
    import cageo_amd as ca
    size = args.size
    dem, stream = ca.synth_dem(size, size, args.seed, 10.0, 2000, NODATA)
    tdem, tst = torch.from_numpy(dem).cuda(), torch.from_numpy(stream).cuda()
    r = ca.filter_radius(60.0, 10.0)
    with ca.Context(size, size, torch.cuda.current_stream()) as ctx:
        ctx.set_filter(ca.GAUSSIAN, r, ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r))
        ctx.mc_setup(tdem, tst, NODATA, NODATA, 2, 0.0, 1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        ctx.run_realizations(mc, 2, args.seed, 0)
    a = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print("probe done", flush=True)


def counter_pass(counter, args, outdir):
    """One rocprofv3 --pmc pass over the probe; returns ({kernel: [values]}, error)."""
    d = os.path.join(outdir, counter)
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__), "--probe", "--size", str(args.size),
           "--seed", str(args.seed)]
    try:
        # rocprofv3 may crash at its own exit after writing complete results
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=120, check=False)
    except subprocess.TimeoutExpired:
        return None, f"{counter}: profiler pass timed out"
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        return None, f"{counter}: no counter file (rc {res.returncode}): {res.stderr[-300:]}"
    vals = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            row = {k.lower(): v for k, v in row.items()}
            if row.get("counter_name") != counter:
                continue
            vals.setdefault(row.get("kernel_name", ""), []).append(float(row["counter_value"]))
    return vals, None


def measure_traffic(args, kernels):
    """HBM bytes per launch of the stage whose launches are `kernels` (read
    and write; the per-launch averages of its kernels summed), each counter
    scaled by what it reads for the calibration kernel (an elementwise
    multiply that reads and writes CALIB_BYTES).  Returns (result or None,
    reason)."""
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    outdir = args.traffic_out or tempfile.mkdtemp(prefix="cageo_pmc_")
    if args.traffic_out:
        os.makedirs(outdir, exist_ok=True)
    try:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            vals, err = counter_pass(counter, args, outdir)
            if vals is None:
                return None, err
            cal = [v for name, v in vals.items() if "AUnaryFunctor" in name]
            per_launch_kib = 0.0
            for kernel in kernels:
                k = [v for name, v in vals.items() if kernel in name]
                if not k or not cal:
                    return None, f"{counter}: {kernel} not found in {sorted(vals)[:8]}"
                per_launch_kib += sum(k[0]) / len(k[0])
            cal_kib = max(cal[0])
            factor = CALIB_BYTES / (cal_kib * 1024.0)
            res[counter] = {"kib_per_launch": per_launch_kib, "calibration_kib": cal_kib,
                            "factor": factor, "bytes": per_launch_kib * 1024.0 * factor}
        return res, None
    except (OSError, ValueError, KeyError) as e:
        return None, f"counter parsing failed: {e}"
    finally:
        if not args.traffic_out:
            shutil.rmtree(outdir, ignore_errors=True)


def batched_leg(torch, ca, args):
    """Batched form: `--batch-tiles` synthetic DEM tiles with sides drawn
    U(256, 2048), per-tile error RMSE U(0.5, 3) m and practical range
    U(30, 200) m (Gaussian filter), `--batch-realizations` each.  Autotuned,
    then timed as one graph replay per realization of all tiles and as
    per-tile blocking calls; the two must give identical counter rasters."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(args.seed + 3)
    n = args.batch_tiles
    shapes = [tuple(int(v) for v in rng.integers(256, 2049, size=2)) for _ in range(n)]
    rmse = rng.uniform(0.5, 3.0, n)
    pr = rng.uniform(30.0, 200.0, n)
    with ThreadPoolExecutor(16) as ex:  # the generator is reentrant and releases the GIL
        dems = list(ex.map(lambda k: ca.synth_dem(shapes[k][1], shapes[k][0], args.seed + 100 + k,
                                                  10.0, 2000, NODATA), range(n)))
    cells = sum(h * w for h, w in shapes)
    dev = [(torch.from_numpy(d).cuda(), torch.from_numpy(s).cuda()) for d, s in dems]
    out = {"tiles": n, "realizations": args.batch_realizations, "cells": cells,
           "sides": "U(256,2048)", "filter": "gaussian"}
    with ca.Batch(shapes) as b:
        for t, (dd, ds) in enumerate(dev):
            b.setup_tile(t, dd, ds, NODATA, NODATA, 2, ca.GAUSSIAN, float(pr[t]), 10.0, 0.0,
                         float(rmse[t]), args.seed + 1000 + t)
        scratch = [torch.zeros(sh, dtype=torch.float32, device="cuda") for sh in shapes]
        t0 = time.perf_counter()
        cfg = b.tune(scratch, reps=1)
        torch.cuda.synchronize()
        out["tune_s"] = round(time.perf_counter() - t0, 2)
        out["tuned_workgroups"] = sorted(set(g for g, _ in cfg))
        out["all_tile_launch"] = b.info()
        res = {}
        for name, mode in [("graph", ca.BATCH_GRAPH), ("per_tile", ca.BATCH_PER_TILE)]:
            mcs = [torch.zeros(sh, dtype=torch.float32, device="cuda") for sh in shapes]
            b.run(mcs, 1, 0, mode)  # warm-up (graph build)
            for m in mcs:
                m.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.run(mcs, args.batch_realizations, 0, mode)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            res[name] = mcs
            out[name] = {"s": round(dt, 3),
                         "tile_realizations_per_s": round(n * args.batch_realizations / dt, 1),
                         "mcells_per_s": round(cells * args.batch_realizations / dt / 1e6, 1)}
        out["identical"] = all(torch.equal(a, c) for a, c in zip(res["graph"], res["per_tile"]))
        out["graph_speedup_vs_per_tile"] = round(out["per_tile"]["s"] / out["graph"]["s"], 3)
    return out


def large_leg(torch, ca, args):
    """Large tile, both growth patterns: the generator's --large-size^2 DEM
    (seed --seed), exponential filter PR 200 m, RMSE 2.5 m.
    - Monte Carlo: --large-realizations realizations in the fast
      convolution mode (the FFT of the torus, within 1e-6 of the exact
      field), in chunks of 50 with one progress line each; its frequency map
      is the reference.  The exact order is timed on 10 realizations beside.
    - Surrogate (exponential weights) through the C ABI: prediction, a
      --large-check realization check on --large-window windows (unbiased
      mean signed error <= 0.01 and RMS error <= 1.2 x 2.46 x 0.02 per
      window), and refinement of the failing windows by a --large-refine
      realization Monte Carlo of the whole DEM.
    Times are wall times with the device synchronised."""
    import numpy as np
    from cageo_amd import surrogate as sg
    log = lambda m: print(f"large: {m}", file=sys.stderr, flush=True)
    size = args.large_size
    t0 = time.perf_counter()
    dem, st = ca.synth_dem(size, size, args.seed, 10.0, 2000, NODATA)
    gen_s = time.perf_counter() - t0
    log(f"synthetic {size}^2 in {gen_s:.1f} s")
    H, W = dem.shape
    ddem, dst = torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda()
    del dem, st
    r = ca.filter_radius(200.0, 10.0)
    taps = ca.create_filter(ca.EXPONENTIAL, 200.0, 10.0, r)
    out = {"dem": f"{W}x{H}", "generator_s": round(gen_s, 1), "filter": "exponential PR 200 m",
           "error_std_m": 2.5, "dem_seed": args.seed}
    seed = args.seed + 11
    n = args.large_realizations
    with ca.Context(W, H, torch.cuda.current_stream()) as ctx:
        ctx.set_filter(ca.EXPONENTIAL, r, taps)
        ctx.mc_setup(ddem, dst, NODATA, NODATA, 2, 0.0, 2.5)
        scratch = torch.zeros((H, W), dtype=torch.float32, device="cuda")
        ctx.run_realizations(scratch, 1, seed, 0)  # exact order, warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        ctx.run_realizations(scratch, 10, seed, 1)
        torch.cuda.synchronize()
        exact_per = (time.perf_counter() - t) / 10
        out["mc_exact_order"] = {"realizations_per_s": round(1 / exact_per, 3),
                                 "est_1000_s": round(1000 * exact_per, 1), "timed": 10}
        log(f"exact order {1 / exact_per:.2f} realizations/s")
        ctx.set_conv_mode(ca.CONV_FAST)
        terms = ctx.conv_terms()
        mc = torch.zeros((H, W), dtype=torch.float32, device="cuda")
        ctx.run_realizations(scratch, 1, seed, 0)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(0, n, 50):
            ctx.run_realizations(mc, min(50, n - k), seed, k)
            torch.cuda.synchronize()
            log(f"Monte Carlo {min(k + 50, n)}/{n} ({time.perf_counter() - t:.1f} s)")
        mc_s = time.perf_counter() - t
        ref = mc / float(n)
        out["mc"] = {"realizations": n,
                     "convolution": "fast: FFT of the torus" if terms == -1 else
                                    f"fast: low-rank, {terms} terms",
                     "note": "the 1000-realization reference runs in the fast convolution "
                             "mode (within 1e-6 of the exact field); the exact order is "
                             "timed beside",
                     "s": round(mc_s, 2), "realizations_per_s": round(n / mc_s, 3),
                     "mcells_per_s": round(W * H * n / mc_s / 1e6, 1)}
        weights = os.path.join(ROOT, "cageo_amd", "weights", "surrogate_exp_b32.bin")
        if os.path.exists(weights):
            win = args.large_window
            bias_tol, rms_tol = sg.check_tolerances(0.02)
            with ca.Surrogate(weights, torch.cuda.current_stream()) as s:
                p = torch.empty((H, W), dtype=torch.float32, device="cuda")
                s.predict(ddem[:1024, :1024].contiguous(), dst[:1024, :1024].contiguous(), NODATA,
                          NODATA, 2.5, 200.0, p[:1024, :1024].contiguous())  # warm-up
                torch.cuda.synchronize()
                t = time.perf_counter()
                s.predict(ddem, dst, NODATA, NODATA, 2.5, 200.0, p)
                torch.cuda.synchronize()
                pred_s = time.perf_counter() - t
            l1_pred = float((p - ref).abs().mean())
            p0 = p.clone()
            t = time.perf_counter()
            acc, bias, rms = ctx.check_probability_map(p, args.large_check, seed + 1, bias_tol,
                                                       rms_tol, scratch, window=win)
            check_s = time.perf_counter() - t
            nbad = int((~acc).sum())
            t = time.perf_counter()
            ctx.refine_probability_map(p, acc, args.large_refine, seed + 2, scratch, window=win)
            torch.cuda.synchronize()
            refine_s = time.perf_counter() - t
            sur_s = pred_s + check_s + refine_s
            log(f"surrogate {sur_s:.2f} s, {nbad} of {acc.size} windows refined")
            dev = (p - ref).abs()
            wl1 = dev.reshape(H // win, win, W // win, win).mean(dim=(1, 3)).cpu().numpy()
            out["surrogate"] = {
                "weights": os.path.relpath(weights, ROOT), "window": win,
                "check_realizations": args.large_check, "bias_tol": bias_tol,
                "rms_tol": round(rms_tol, 5), "predict_s": round(pred_s, 3),
                "check_s": round(check_s, 3), "windows": int(acc.size),
                "windows_refined": nbad, "refine_realizations": args.large_refine,
                "refine_s": round(refine_s, 2), "total_s": round(sur_s, 2),
                "speedup_vs_mc": round(mc_s / sur_s, 1),
                "l1_prediction": round(l1_pred, 5), "l1_final": round(float(dev.mean()), 5),
                "max_window_l1_final": round(float(wl1.max()), 5),
                "max_dev_final": round(float(dev.max()), 4),
                "check_rms_max": round(float(rms.max()), 5),
                "check_bias_absmax": round(float(np.abs(bias).max()), 5)}
            # the same check and refinement at a stricter target (mean
            # absolute error 0.01), so the refinement path runs in every
            # bench: windows refined, and their error against the reference
            # before and after
            bt2, rt2 = sg.check_tolerances(0.01)
            q = p0.clone()
            t = time.perf_counter()
            acc2, _, _ = ctx.check_probability_map(q, args.large_check, seed + 3, bt2, rt2,
                                                   scratch, window=win)
            ctx.refine_probability_map(q, acc2, args.large_refine, seed + 4, scratch, window=win)
            torch.cuda.synchronize()
            strict_s = time.perf_counter() - t
            bad = ~acc2.reshape(H // win, W // win)
            w_before = (p0 - ref).abs().reshape(H // win, win, W // win, win).mean(dim=(1, 3))
            w_after = (q - ref).abs().reshape(H // win, win, W // win, win).mean(dim=(1, 3))
            nb2 = int(bad.sum())
            bad_t = torch.from_numpy(bad).cuda()
            out["surrogate"]["strict_target"] = {
                "l1_target": 0.01, "bias_tol": bt2, "rms_tol": round(rt2, 5),
                "windows_refined": nb2, "s": round(strict_s, 2),
                "refined_windows_l1_before": round(float(w_before[bad_t].mean()), 5) if nb2 else None,
                "refined_windows_l1_after": round(float(w_after[bad_t].mean()), 5) if nb2 else None,
                "l1_final": round(float((q - ref).abs().mean()), 5)}
            log(f"strict target: {nb2} windows refined in {strict_s:.1f} s")
    return out


def surrogate_leg(torch, ca, dem, stream, tdem, tst, args):
    """The packaged network (C-ABI inference from the resident DEM) with its
    Monte Carlo check, against an N-realization Monte Carlo of the same DEM
    and error model: mean absolute and tail per-cell deviation of the border
    probability, and the wall time of each (device work bracketed by
    synchronisations; both with the context already set up)."""
    from cageo_amd import surrogate as sg
    n = args.surrogate_realizations
    size = dem.shape[0]
    sh = torch.cuda.current_stream()
    r = ca.filter_radius(60.0, 10.0)
    with ca.Context(size, size, sh) as ctx, ca.Surrogate(sg.DEFAULT_WEIGHTS, sh) as s:
        ctx.set_filter(ca.GAUSSIAN, r, ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r))
        ctx.mc_setup(tdem, tst, NODATA, NODATA, 2, 0.0, 1.8)
        cnt = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        ctx.run_realizations(cnt, 2, args.seed, 0)  # warm-up
        cnt.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run_realizations(cnt, n, args.seed, 0)
        torch.cuda.synchronize()
        mc_s = time.perf_counter() - t0
        p_mc = cnt / float(n)
        q = torch.empty((size, size), dtype=torch.float32, device="cuda")
        s.predict(tdem, tst, NODATA, NODATA, 1.8, 60.0, q)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.predict(tdem, tst, NODATA, NODATA, 1.8, 60.0, q)
        torch.cuda.synchronize()
        pred_s = time.perf_counter() - t0
        out, info = sg.checked_prediction(torch, None, dem, stream, 1.8, 60.0, args.seed + 7,
                                          fallback_realizations=n, prediction=q, ctx=ctx)
        sur_s = pred_s + info["check_s"] + info.get("fallback_s", 0.0)
        # the prediction alone, timed with events on its stream
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(sh)
        for _ in range(reps):
            s.predict(tdem, tst, NODATA, NODATA, 1.8, 60.0, q)
        e1.record(sh)
        torch.cuda.synchronize()
        fwd_ms = e0.elapsed_time(e1) / reps
    # the prediction's first layer reads the 4 input planes only
    flops = sg.forward_flops(size, size, in_channels=sg.IN_CHANNELS)
    tflops = flops / (fwd_ms * 1e-3) / 1e12
    dev = (out - p_mc).abs()
    border = p_mc > 0
    return {
        "mc_realizations": n, "mc_s": round(mc_s, 4), "surrogate_s": round(sur_s, 4),
        "predict_s": round(pred_s, 4), "speedup_vs_mc": round(mc_s / sur_s, 1),
        "l1": round(float(dev.mean()), 5),
        "l1_border_cells": round(float(dev[border].mean()), 5) if bool(border.any()) else 0.0,
        "p999_dev": round(float(torch.quantile(dev.flatten()[::16].float(), 0.999)), 4),
        "max_dev": round(float(dev.max()), 4),
        "check": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in info.items()},
        "weights": os.path.relpath(sg.DEFAULT_WEIGHTS, ROOT),
        "predict_ms": round(fwd_ms, 3),
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / BF16_PEAK_TFLOPS, 5),
                     "flops_per_forward": flops,
                     "note": "BF16 MFMA convolutions of one 4096x4096 prediction, the first "
                             "layer counted at its 4 input channels (input statistics and "
                             "crop included in the time)"},
    }


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the Monte Carlo "
              "realizations are independent and would shard by realization index, but no "
              "multi-GPU run is part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.probe:
        probe(args)
        return

    import torch
    import cageo_amd as ca

    oracle_ok = os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so"))
    size = args.size
    cells = size * size
    dem, stream = ca.synth_dem(size, size, args.seed, 10.0, 2000, NODATA)
    torch.cuda.init()
    tdem = torch.from_numpy(dem).cuda()
    tst = torch.from_numpy(stream).cuda()
    stream_h = torch.cuda.current_stream()

    def run(model, pr, steps, warmup, stage_times, conv_mode=ca.CONV_EXACT):
        r = ca.filter_radius(pr, 10.0)
        taps = ca.create_filter(model, pr, 10.0, r)
        ctx = ca.Context(size, size, stream_h)
        ctx.set_filter(model, r, taps)
        ctx.set_conv_mode(conv_mode)
        run.terms = ctx.conv_terms()
        ctx.mc_setup(tdem, tst, NODATA, NODATA, 2, 0.0, 1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        ctx.stage_times(stage_times)
        ctx.run_realizations(mc, warmup, args.seed, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run_realizations(mc, steps, args.seed, warmup)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stages = None
        if stage_times:
            ms = ctx.stage_times(False) / steps
            stages = dict(zip(STAGE_NAMES, [round(float(v), 4) for v in ms]))
        total = float(mc.sum().item())
        ctx.close()
        return dt, stages, total

    log(f"Gaussian PR 60 m, {args.steps} realizations")
    dt, stages, total = run(ca.GAUSSIAN, 60.0, args.steps, args.warmup,
                            not args.no_stage_times)
    real_s = args.steps / dt
    line = {
        "metric": METRIC,
        "value": round(real_s, 3),
        "unit": "realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"MC drainage realizations, {size}x{size} synthetic DEM, 10 m cells, "
                        "Gaussian filter PR 60 m, error std 1.8 m",
            "dem": f"{size}x{size}", "filter": "gaussian", "practical_range_m": 60,
            "error_std_m": 1.8, "dem_seed": args.seed, "parallelism": "single GPU",
        },
        "mcells_per_s": round(cells * real_s / 1e6, 2),
        "border_count_sum": total,
    }
    if stages:
        line["stage_ms"] = stages
        key = max(STAGE_KERNEL, key=lambda k: stages[k])
        kernels = STAGE_KERNEL[key]
        bytes_per_launch = STAGE_BYTES[key] * cells
        achieved = bytes_per_launch / (stages[key] * 1e-3) / 1e9
        log("counter passes" if not args.no_traffic else "counter passes skipped")
        traffic, why = (measure_traffic(args, kernels) if not args.no_traffic
                        else (None, "not requested"))
        if traffic is None and not args.no_traffic:
            print(f"bench.py: traffic not measured: {why}", file=sys.stderr)
        line["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": (round(traffic["FETCH_SIZE"]["bytes"] + traffic["WRITE_SIZE"]["bytes"])
                        if traffic else None),
            "kernel": " + ".join(kernels), "launch_ms": stages[key],
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_per_cell": STAGE_BYTES[key],
        }
        if traffic:
            line["roofline"]["traffic_detail"] = {
                k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in traffic.items()}
        # per stage and end to end: algorithmic bytes of every stage per
        # realization over the measured time (stage events; whole step)
        line["roofline_stages"] = {
            k: {"gb_s": round(STAGE_BYTES[k] * cells / (stages[k] * 1e-3) / 1e9, 1),
                "frac": round(STAGE_BYTES[k] * cells / (stages[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
            for k in STAGE_NAMES if stages.get(k)}
    e2e_bytes = sum(STAGE_BYTES.values()) * cells
    e2e = e2e_bytes / (dt / args.steps) / 1e9
    line["roofline_end_to_end"] = {
        "bound": "hbm", "achieved": round(e2e, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(e2e / HBM_PEAK_GBS, 5), "bytes_per_realization": e2e_bytes,
        "note": "80 B/cell per realization (one pass over every stage's boundary I/O at the "
                "reference's data types, fused stages counted unfused) over the whole step"}
    if args.exp_steps > 0:
        log(f"exponential PR 200 m, {args.exp_steps} realizations, exact and fast")
        dte, st_e, _ = run(ca.EXPONENTIAL, 200.0, args.exp_steps, 1, not args.no_stage_times)
        line["exponential_pr200"] = {
            "realizations_per_s": round(args.exp_steps / dte, 3),
            "mcells_per_s": round(cells * args.exp_steps / dte / 1e6, 2),
            "steps": args.exp_steps, "stage_ms": st_e,
        }
        if st_e:
            conv_ms = st_e["convolution"]
            flops = 2.0 * (2 * ca.filter_radius(200.0, 10.0) + 1) ** 2 * cells
            tf = flops / conv_ms / 1e9
            line["exponential_pr200"]["convolution_roofline"] = {
                "bound": "fp32-vector", "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 4),
                "note": "k_conv_2d; no FMA (exact reference order)"}
        dtf, st_f, _ = run(ca.EXPONENTIAL, 200.0, args.exp_steps, 1, not args.no_stage_times,
                           ca.CONV_FAST)
        fast = {"realizations_per_s": round(args.exp_steps / dtf, 3),
                "mcells_per_s": round(cells * args.exp_steps / dtf / 1e6, 2),
                "form": "fft" if run.terms < 0 else "low-rank", "stage_ms": st_f,
                "note": "periodic correlation through FFTs of the W x H torus (power-of-two "
                        "sides), within 1e-5 of the exact convolution (not bit-exact)"
                        if run.terms < 0 else
                        "low-rank separable form of the filter with fused multiply-adds, within "
                        "1e-5 of the exact convolution (not bit-exact)"}
        if st_f and run.terms < 0:
            b = 24 * cells  # three passes, each reading and writing 4 B per cell
            gbs = b / (st_f["convolution"] * 1e-3) / 1e9
            fast["convolution_roofline"] = {
                "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 24,
                "note": "k_fft_rows_fwd + k_fft_cols + k_fft_rows_inv: row FFTs to the half "
                        "spectrum, column FFT x filter spectrum x inverse column FFT, inverse "
                        "row FFTs; 8 B per cell each"}
        elif st_f:
            K, L = run.terms, 2 * ca.filter_radius(200.0, 10.0) + 1
            flops = 2.0 * 2 * K * L * cells  # column and row passes, one FMA = 2 FLOP
            tf = flops / st_f["convolution"] / 1e9
            fast["convolution_roofline"] = {
                "bound": "fp32-vector", "achieved": round(tf, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tf / FP32_PEAK_TFLOPS, 4),
                "note": "k_lr_cols + k_lr_rows, 2 x K x (2r+1) FMA per cell"}
        line["exponential_pr200"]["fast_mode"] = fast
    if args.batch_tiles > 0:
        log(f"batched form, {args.batch_tiles} tiles")
        line["batched"] = batched_leg(torch, ca, args)
    if args.surrogate_realizations > 0:
        log("surrogate vs Monte Carlo")
        line["surrogate"] = surrogate_leg(torch, ca, dem, stream, tdem, tst, args)
    if args.large_size > 0:
        line["large_tile"] = large_leg(torch, ca, args)
    if args.cpu_sample > 0 and oracle_ok:
        log(f"CPU baseline, {args.cpu_sample} realizations")
        cps, secs = cpu_baseline(size, args.cpu_sample, args.seed)
        line["cpu_baseline"] = {
            "value": round(cps, 5), "unit": "realizations/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_sample} realizations of the same {size}x{size} Gaussian "
                      f"workload on the CPU oracle, OMP_NUM_THREADS=1 ({secs:.1f} s)",
        }
        line["speedup_vs_cpu_1core"] = round(real_s / cps, 1)
        nt = cpu_threads()
        cpa, secs_a = cpu_baseline(size, args.cpu_sample, args.seed, nt)
        line["cpu_all_cores"] = {
            "value": round(cpa, 5), "unit": "realizations/s", "cores": nt,
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{args.cpu_sample} realizations, OMP_NUM_THREADS={nt} ({secs_a:.1f} s); "
                      "the oracle's OpenMP loops follow the reference's (pit fill and tracing "
                      "are sequential there too)",
            "gpu_speedup": round(real_s / cpa, 1)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
