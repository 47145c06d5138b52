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
one GPU.  `roofline` covers the dominant kernel (most device time per
realization), measured with HIP events inside the library on its stream.
`cpu_baseline` times the CPU oracle (a restatement of the reference's OpenMP
CPU path) on a bounded sample, in a subprocess pinned to one thread.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "realizations/s + Mcells/s at 4096² DEM; surrogate vs MC time @ L1≤0.02"
HBM_PEAK_GBS = 8000.0
NODATA = -9999.0

# Algorithmic bytes per cell per realization of each stage, at the
# reference's data types (one pass over stage-boundary I/O; iterations of the
# fixed-point stages are not counted).
STAGE_BYTES = {
    "random_field": 4, "convolution_gaussian": 16, "convolution_exponential": 8,
    "realistic_burn": 24, "pit_fill": 8, "flow_routing": 6, "flow_tracing": 10,
    "border": 12,
}
STAGE_NAMES = ["random_field", "convolution", "realistic_burn", "pit_fill", "flow_routing",
               "flow_tracing", "border"]


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=4096)
    p.add_argument("--exp-steps", type=int, default=20,
                   help="realizations timed with the exponential filter (0 = skip)")
    p.add_argument("--seed", type=int, default=20240601)
    p.add_argument("--cpu-sample", type=int, default=2,
                   help="realizations timed on the CPU oracle (0 = skip)")
    p.add_argument("--no-stage-times", action="store_true")
    return p.parse_args()


def cpu_baseline(size, sample, seed):
    """Time the CPU oracle on `sample` realizations (1 thread) in a child."""
    code = f"""
import os, sys, time, json
sys.path.insert(0, {ROOT!r})
import numpy as np, oracle, cageo_amd as ca
dem, st = ca.synth_dem({size}, {size}, {seed}, 10.0, 2000, {NODATA})
r = oracle.filter_radius(60.0, 10.0)
f = oracle.create_filter(oracle.GAUSSIAN, 60.0, 10.0, r)
mc = np.zeros(dem.shape, np.float32)
t = time.perf_counter()
for k in range({sample}):
    st_ = oracle.lib().or_realization(dem, st, {size}, {size}, f, r, 1, 0.0, 1.8, {seed}, k,
                                      {NODATA}, {NODATA}, 2, mc)
    assert st_ == 0
dt = time.perf_counter() - t
print(json.dumps({{"seconds": dt, "n": {sample}}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         check=True)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    return res["n"] / res["seconds"], res["seconds"]


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the Monte Carlo "
              "realizations are independent and would shard by realization index, but no "
              "multi-GPU run is part of this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import cageo_amd as ca

    oracle_ok = os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so"))
    size = args.size
    dem, stream = ca.synth_dem(size, size, args.seed, 10.0, 2000, NODATA)
    torch.cuda.init()
    tdem = torch.from_numpy(dem).cuda()
    tst = torch.from_numpy(stream).cuda()
    stream_h = torch.cuda.current_stream()

    def run(model, pr, steps, warmup, stage_times):
        r = ca.filter_radius(pr, 10.0)
        taps = ca.create_filter(model, pr, 10.0, r)
        ctx = ca.Context(size, size, stream_h)
        ctx.set_filter(model, r, taps)
        ctx.mc_setup(tdem, tst, NODATA, NODATA, 2, 0.0, 1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        ctx.run_realizations(mc, warmup, args.seed, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.run_realizations(mc, steps, args.seed, warmup)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stages = None
        if stage_times:
            ctx.stage_times(True)
            n_t = max(1, min(steps, 10))
            ctx.run_realizations(mc, n_t, args.seed, warmup + steps)
            ms = ctx.stage_times(False) / n_t
            stages = dict(zip(STAGE_NAMES, [round(float(v), 4) for v in ms]))
        total = float(mc.sum().item())
        ctx.close()
        return dt, stages, total

    dt, stages, total = run(ca.GAUSSIAN, 60.0, args.steps, args.warmup,
                            not args.no_stage_times)
    real_s = args.steps / dt
    cells = size * size
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
        # dominant stage by device time; bytes per realization at the
        # reference's data types (STAGE_BYTES)
        key = max(stages, key=lambda k: stages[k])
        per_cell = STAGE_BYTES["convolution_gaussian" if key == "convolution" else key]
        achieved = per_cell * cells / (stages[key] * 1e-3) / 1e9
        line["roofline"] = {
            "bound": "hbm", "kernel": key, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "bytes_per_cell": per_cell,
        }
    if args.exp_steps > 0:
        dte, st_e, _ = run(ca.EXPONENTIAL, 200.0, args.exp_steps, 1, not args.no_stage_times)
        line["exponential_pr200"] = {
            "realizations_per_s": round(args.exp_steps / dte, 3),
            "mcells_per_s": round(cells * args.exp_steps / dte / 1e6, 2),
            "steps": args.exp_steps, "stage_ms": st_e,
        }
        if st_e:
            conv_ms = st_e["convolution"]
            flops = 2.0 * (2 * ca.filter_radius(200.0, 10.0) + 1) ** 2 * cells
            line["exponential_pr200"]["convolution_tflops"] = round(flops / conv_ms / 1e9, 2)
    if args.cpu_sample > 0 and oracle_ok:
        cps, secs = cpu_baseline(size, args.cpu_sample, args.seed)
        line["cpu_baseline"] = {
            "value": round(cps, 5), "unit": "realizations/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_sample} realizations of the same {size}x{size} Gaussian "
                      f"workload on the CPU oracle, OMP_NUM_THREADS=1 ({secs:.1f} s)",
        }
        line["speedup_vs_cpu_1core"] = round(real_s / cps, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
