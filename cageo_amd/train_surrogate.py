"""Train the border-probability surrogate on Monte Carlo maps and judge it.

    python -m cageo_amd.train_surrogate --out weights.bin [--report report.json]
    python -m cageo_amd.train_surrogate --load default [--report report.json]

With --load only step 3 runs, on the given (or the packaged) weights.

1. Training set: `--n-dems` synthetic DEMs of `--size`^2 (the evaluation
   generator, other seeds), each with RMSE ~ U(0.5, 3) m and practical range
   ~ U(30, `--pr-max`) m, labelled by `--realizations` Monte Carlo
   realizations of the whole DEM, cut into `--crop`^2 tiles; inputs and
   labels stay on the device (the defaults: 320 DEMs of 4096^2, 20480 tiles
   of 512^2).
2. Training: `--steps` Adam steps of `--batch` tiles through the C ABI
   (cageo_surrogate_train) on the library's BF16 MFMA kernels.
3. Judgement: on the `--eval-size`^2 DEM (seed `--eval-seed`, RMSE
   `--eval-rmse`, practical range `--eval-range`) the prediction is compared
   with a `--eval-realizations` Monte Carlo: mean absolute and maximum
   per-cell deviation, and the wall time of the checked prediction
   (inference + `--check-realizations` Monte Carlo check) against the time of
   that Monte Carlo.
Progress goes to stdout; the last line is the JSON report.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="weights file (.bin) to write")
    ap.add_argument("--load", default=None,
                    help="judge these weights instead of training ('default': the packaged ones)")
    ap.add_argument("--report", default=None, help="also write the JSON report here")
    ap.add_argument("--n-dems", type=int, default=320)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--realizations", type=int, default=200)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--eval-size", type=int, default=4096)
    ap.add_argument("--eval-seed", type=int, default=20240601)
    ap.add_argument("--eval-rmse", type=float, default=1.8)
    ap.add_argument("--eval-range", type=float, default=60.0)
    ap.add_argument("--eval-realizations", type=int, default=1000)
    ap.add_argument("--check-realizations", type=int, default=4)
    ap.add_argument("--filter", choices=["Gaussian", "Exponential"], default="Gaussian",
                    help="error-field filter of the Monte Carlo labels and the judgement")
    ap.add_argument("--pr-max", type=float, default=200.0,
                    help="training practical ranges are drawn U(30, pr-max) m")
    ap.add_argument("--fast-convolution", action="store_true",
                    help="label with the fast convolution mode (FFT of the torus on power-of-two "
                         "sides; within 1e-5 of the exact field)")
    a = ap.parse_args(argv)

    import torch
    import cageo_amd as ca
    from cageo_amd import surrogate as sg

    def log(msg):
        print(msg, flush=True)

    model = ca.GAUSSIAN if a.filter == "Gaussian" else ca.EXPONENTIAL

    if a.load:
        net = sg.load_net(torch, None if a.load == "default" else a.load)
        hist, train_s, label_s = [], 0.0, 0.0
        log(f"weights <- {a.load}")
    else:
        if not a.out:
            ap.error("--out is required when training")
        t0 = time.perf_counter()
        x4, y, meta = sg.training_set(torch, a.n_dems, a.size, a.crop, a.realizations, a.seed,
                                      model=model, pr_range=(30.0, a.pr_max), log=log,
                                      fast_convolution=a.fast_convolution)
        torch.cuda.synchronize()
        label_s = time.perf_counter() - t0
        ntiles = int(x4.shape[0])
        log(f"training set: {ntiles} tiles of {a.crop}^2 from {a.n_dems} DEMs of "
            f"{a.size}^2, {a.realizations} realizations each ({label_s:.1f} s)")
        t1 = time.perf_counter()
        tr, hist = sg.train(torch, x4, y, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed,
                            log=log)
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t1
        tr.save(a.out)
        tr.close()
        log(f"trained {a.steps} steps in {train_s:.1f} s, weights -> {a.out}")
        del x4, y
        net = sg.load_net(torch, a.out)

    dem, st = ca.synth_dem(a.eval_size, a.eval_size, a.eval_seed, 10.0, 2000, -9999.0)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    p_mc = sg.mc_frequency(torch, [dem], [st], [a.eval_rmse], [a.eval_range],
                           a.eval_realizations, a.eval_seed, on_device=True, model=model)[0]
    torch.cuda.synchronize()
    mc_s = time.perf_counter() - t2
    log(f"judge: {a.eval_realizations}-realization Monte Carlo in {mc_s:.2f} s")
    sg.checked_prediction(torch, net, dem[:512, :512], st[:512, :512], a.eval_rmse, a.eval_range,
                          1, check_realizations=2, fallback_realizations=2, model=model)  # warm-up
    q, info = sg.checked_prediction(torch, net, dem, st, a.eval_rmse, a.eval_range,
                                    a.eval_seed + 7, check_realizations=a.check_realizations,
                                    fallback_realizations=a.eval_realizations, model=model)
    surrogate_s = info["predict_s"] + info["check_s"] + info.get("fallback_s", 0.0)
    dev = (q - p_mc).abs().cpu().numpy()
    p_mc = p_mc.cpu().numpy()
    report = {
        "weights": a.load or a.out,
        "train_tiles": int(len(hist) and a.n_dems * (a.size // a.crop) ** 2),
        "train_dems": f"{a.n_dems} x {a.size}^2" if hist else None,
        "label_s": round(label_s, 1) if hist else None,
        "train_batch": a.batch if hist else None,
        "label_convolution": ("fast" if a.fast_convolution else "exact") if hist else None,
        "train_realizations": a.realizations, "train_steps": len(hist), "train_s": round(train_s, 2),
        "final_loss": float(np.mean(hist[-100:])) if hist else None,
        "eval_dem": f"{a.eval_size}x{a.eval_size} seed {a.eval_seed}",
        "eval_rmse_m": a.eval_rmse, "eval_practical_range_m": a.eval_range, "filter": a.filter,
        "mc_realizations": a.eval_realizations, "mc_s": round(mc_s, 3),
        "l1": float(dev.mean()), "max_dev": float(dev.max()),
        "p999_dev": float(np.quantile(dev, 0.999)),
        "l1_border_cells": float(dev[p_mc > 0].mean()) if (p_mc > 0).any() else 0.0,
        "check": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in info.items()},
        "surrogate_s": round(surrogate_s, 4),
        "speedup_vs_mc": round(mc_s / surrogate_s, 1),
    }
    line = json.dumps(report)
    if a.report:
        with open(a.report, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
