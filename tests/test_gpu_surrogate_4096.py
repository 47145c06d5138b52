"""The learned surrogate at 4096 x 4096, judged by the Monte Carlo.

- The packaged network (C-ABI inference) on the benchmark DEM and on two
  DEMs outside its training set (synthetic 4096^2, seeds 20240601, 20240611,
  20240621; Gaussian PR 60 m, RMSE 1.8 m): the Monte Carlo check accepts the
  map, its mean absolute deviation from a 1000-realization Monte Carlo is at
  most 0.02, and on the benchmark DEM prediction + check take at most 1/50 of
  that Monte Carlo's wall time (both with the context already set up).
- The check's power: maps at a true mean absolute deviation of 0.03 and 0.05
  from the 1000-realization map, made from the prediction in six ways (its
  error amplified, blurred, biased up and down, scaled up and down), are all
  rejected, while the prediction itself is accepted in the same setting.

Each case prints its name before it touches the GPU.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE = 4096
SEEDS = [20240601, 20240611, 20240621]


def announce(*what):
    print("CASE", *what, flush=True)


@pytest.fixture(scope="module")
def env(torch_cuda):
    import cageo_amd as ca
    from cageo_amd import surrogate as sg
    ca.lib()
    return torch_cuda, ca, sg


def _setup(torch, ca, seed):
    dem, st = ca.synth_dem(SIZE, SIZE, seed, 10.0, 2000, -9999.0)
    ddem, dst = torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda()
    r = ca.filter_radius(60.0, 10.0)
    ctx = ca.Context(SIZE, SIZE, torch.cuda.current_stream())
    ctx.set_filter(ca.GAUSSIAN, r, ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r))
    ctx.mc_setup(ddem, dst, -9999.0, -9999.0, 2, 0.0, 1.8)
    return dem, st, ddem, dst, ctx


def _mc(torch, ctx, n, seed):
    cnt = torch.zeros((SIZE, SIZE), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    ctx.run_realizations(cnt, n, seed, 0)
    torch.cuda.synchronize()
    return cnt / float(n), time.perf_counter() - t


def _predict(torch, ca, sg, ddem, dst, s):
    q = torch.empty((SIZE, SIZE), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    s.predict(ddem, dst, -9999.0, -9999.0, 1.8, 60.0, q)
    torch.cuda.synchronize()
    return q, time.perf_counter() - t


@pytest.fixture(scope="module")
def bench_dem(env):
    torch, ca, sg = env
    dem, st, ddem, dst, ctx = _setup(torch, ca, SEEDS[0])
    announce("surrogate 4096^2: 1000-realization Monte Carlo, seed", SEEDS[0])
    _mc(torch, ctx, 2, 1)  # warm-up
    p, mc_s = _mc(torch, ctx, 1000, SEEDS[0] + 5)
    yield dem, st, ddem, dst, ctx, p, mc_s
    ctx.close()


def test_surrogate_4096_vs_1000_realizations(env, bench_dem):
    torch, ca, sg = env
    dem, st, ddem, dst, ctx, p, mc_s = bench_dem
    announce("surrogate 4096^2 gain on the benchmark DEM")
    with ca.Surrogate(sg.DEFAULT_WEIGHTS, torch.cuda.current_stream()) as s:
        q, pred_s = _predict(torch, ca, sg, ddem, dst, s)  # warm-up of the full size
        q, pred_s = _predict(torch, ca, sg, ddem, dst, s)
    out, info = sg.checked_prediction(torch, None, dem, st, 1.8, 60.0, 77, prediction=q, ctx=ctx)
    sur_s = pred_s + info["check_s"]
    l1 = float((q - p).abs().mean())
    print(f"l1 {l1:.5f} MC {mc_s:.3f} s surrogate {sur_s:.4f} s ({mc_s / sur_s:.0f}x) {info}",
          flush=True)
    assert info["accepted"], info
    assert l1 <= 0.02
    assert mc_s >= 50.0 * sur_s, (mc_s, sur_s)


@pytest.mark.parametrize("seed", SEEDS[1:])
def test_surrogate_4096_held_out_dems(env, seed):
    torch, ca, sg = env
    announce("surrogate 4096^2 on a held-out DEM, seed", seed)
    dem, st, ddem, dst, ctx = _setup(torch, ca, seed)
    try:
        p, _ = _mc(torch, ctx, 1000, seed + 5)
        with ca.Surrogate(sg.DEFAULT_WEIGHTS, torch.cuda.current_stream()) as s:
            q, _ = _predict(torch, ca, sg, ddem, dst, s)
        _, info = sg.checked_prediction(torch, None, dem, st, 1.8, 60.0, 78, prediction=q,
                                        ctx=ctx)
        l1 = float((q - p).abs().mean())
        print(f"l1 {l1:.5f} {info}", flush=True)
        assert info["accepted"], info
        assert l1 <= 0.02
    finally:
        ctx.close()


def _gauss_blur(torch, q, sigma):
    r = max(1, int(3 * sigma + 0.5))
    x = torch.arange(-r, r + 1, device="cuda", dtype=torch.float32)
    k = torch.exp(-x * x / (2 * sigma * sigma))
    k = k / k.sum()
    t = q[None, None]
    t = torch.nn.functional.conv2d(torch.nn.functional.pad(t, (r, r, 0, 0), mode="replicate"),
                                   k.view(1, 1, 1, -1))
    t = torch.nn.functional.conv2d(torch.nn.functional.pad(t, (0, 0, r, r), mode="replicate"),
                                   k.view(1, 1, -1, 1))
    return t[0, 0]


def _at_l1(torch, fam, p, lo, hi, target):
    """fam(param) with mean |fam - p| = target (bisection; None if out of reach)."""
    if float((fam(hi) - p).abs().mean()) < target:
        return None
    for _ in range(24):
        mid = 0.5 * (lo + hi)
        if float((fam(mid) - p).abs().mean()) < target:
            lo = mid
        else:
            hi = mid
    return fam(hi)


def test_check_rejects_maps_at_l1_003_and_005(env, bench_dem):
    torch, ca, sg = env
    dem, st, ddem, dst, ctx, p, _ = bench_dem
    with ca.Surrogate(sg.DEFAULT_WEIGHTS, torch.cuda.current_stream()) as s:
        q, _ = _predict(torch, ca, sg, ddem, dst, s)
    announce("check power: the prediction itself")
    _, info = sg.checked_prediction(torch, None, dem, st, 1.8, 60.0, 91, prediction=q, ctx=ctx,
                                    fallback_realizations=2)
    assert info["accepted"], info
    fams = {
        "error amplified": (lambda a: (p + a * (q - p)).clamp(0, 1), 1.0, 20.0),
        "blurred": (lambda sgm: _gauss_blur(torch, q, sgm).clamp(0, 1), 0.05, 12.0),
        "biased up": (lambda b: (q + b).clamp(0, 1), 0.0, 0.5),
        "biased down": (lambda b: (q - b).clamp(0, 1), 0.0, 0.5),
        "scaled up": (lambda a: (q * a).clamp(0, 1), 1.0, 4.0),
        "scaled down": (lambda a: (q * (1.0 - a)).clamp(0, 1), 0.0, 1.0),
    }
    tried = 0
    for target in (0.03, 0.05):
        for name, (fam, lo, hi) in fams.items():
            m = _at_l1(torch, fam, p, lo, hi, target)
            if m is None:
                print(f"{name}: L1 {target} out of reach", flush=True)
                continue
            tried += 1
            l1 = float((m - p).abs().mean())
            announce("check power:", name, "true L1", round(l1, 4))
            _, info = sg.checked_prediction(torch, None, dem, st, 1.8, 60.0, 92, prediction=m,
                                            ctx=ctx, fallback_realizations=2)
            print(f"  {info}", flush=True)
            assert not info["accepted"], (name, l1, info)
    assert tried >= 10
