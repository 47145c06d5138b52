"""Surrogate kernels and network on the GPU against the float64 oracle.

Inputs and weights are rounded to bfloat16 first, so the only differences
left are the kernels' FP32 accumulation order and, for bf16 outputs, the
output rounding (relative 2^-8).  Tolerances are stated per check.  Each
case prints its name before it touches the GPU.
"""
import ctypes

import numpy as np
import pytest

from oracle import unet

pytestmark = pytest.mark.gpu


def announce(*what):
    print("CASE", *what, flush=True)


def to_bf16_bits(a):
    from cageo_amd.surrogate import bf16_round
    u = bf16_round(np.asarray(a, np.float32)).view(np.uint32) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b).astype(np.uint32) << 16).view(np.float32)


@pytest.fixture(scope="module")
def env(torch_cuda):
    import cageo_amd.surrogate as sg
    torch = torch_cuda
    L = sg._nn()
    return torch, sg, L


def dev_bits(torch, a):
    return torch.from_numpy(to_bf16_bits(a).view(np.int16)).cuda()


def host_bits(t):
    return from_bf16_bits(t.cpu().numpy().view(np.uint16))


CONV_CASES = [  # n, h, w, c1, up1, c2, cout, relu
    (2, 16, 20, 32, 0, 0, 64, 1),
    (1, 9, 7, 64, 1, 32, 32, 0),
    (1, 33, 17, 32, 0, 0, 16, 1),
    (3, 8, 8, 96, 0, 0, 128, 1),
    # widths of 64 and heights of 4 take the LDS-staged kernel
    (2, 8, 128, 64, 1, 32, 64, 1),
    (1, 12, 64, 32, 0, 0, 32, 0),
    (1, 4, 192, 128, 0, 0, 96, 1),
]


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv_forward_and_gradients(env, case):
    torch, sg, L = env
    n, h, w, c1, up1, c2, cout, relu = case
    announce("nn conv", case)
    r = np.random.default_rng(sum(case))
    h1, w1 = ((h + 1) // 2, (w + 1) // 2) if up1 else (h, w)
    x1 = sg.bf16_round(r.normal(size=(n, h1, w1, c1)).astype(np.float32))
    x2 = sg.bf16_round(r.normal(size=(n, h, w, c2)).astype(np.float32)) if c2 else None
    wt = sg.bf16_round((r.normal(size=(9, cout, c1 + c2)) * 0.1).astype(np.float32))
    b = r.normal(size=cout).astype(np.float32)
    xin = np.repeat(np.repeat(x1, 2, 1), 2, 2)[:, :h, :w] if up1 else x1
    if c2:
        xin = np.concatenate([xin, x2], axis=-1)
    ref = unet.conv3x3(xin.astype(np.float64), wt.astype(np.float64), b)
    if relu:
        ref = np.maximum(ref, 0)
    d1, dw_ = dev_bits(torch, x1), dev_bits(torch, wt)
    d2 = dev_bits(torch, x2) if c2 else None
    db = torch.from_numpy(b).cuda()
    out = torch.empty((n, h, w, cout), dtype=torch.int16, device="cuda")
    outf = torch.empty((n, h, w, cout), dtype=torch.float32, device="cuda")
    p = sg._p
    assert L.cageo_nn_conv3x3(None, p(d1), c1, up1, p(d2), c2, p(dw_), p(db), p(out), p(outf), n,
                              h, w, cout, relu) == 0
    torch.cuda.synchronize()
    scale = np.abs(ref).max()
    # fp32 result: accumulation order only
    assert np.abs(outf.cpu().numpy() - ref).max() <= 1e-5 * scale * (c1 + c2) ** 0.5
    # bf16 result: plus one rounding of the output (relative 2^-8)
    assert np.abs(host_bits(out) - ref).max() <= 2 ** -8 * scale + 1e-5 * scale * (c1 + c2) ** 0.5

    # weight and input gradients for an output gradient g (bf16)
    g = sg.bf16_round(r.normal(size=(n, h, w, cout)).astype(np.float32))
    dx_ref, dw_ref, _ = unet.conv3x3_back(xin.astype(np.float64), wt.astype(np.float64),
                                          g.astype(np.float64))
    dg = dev_bits(torch, g)
    dwt = torch.zeros((9, cout, c1 + c2), dtype=torch.float32, device="cuda")
    if cout % 32:  # the weight gradient works on 32-channel blocks
        assert L.cageo_nn_conv3x3_wgrad(None, p(d1), c1, up1, p(d2), c2, p(dg), p(dwt), n, h, w,
                                        cout) == -1
        return
    wflip = torch.empty((9, c1 + c2, cout), dtype=torch.int16, device="cuda")
    assert L.cageo_nn_flip_transpose(None, p(dw_), p(wflip), cout, c1 + c2) == 0
    assert L.cageo_nn_conv3x3_wgrad(None, p(d1), c1, up1, p(d2), c2, p(dg), p(dwt), n, h, w,
                                    cout) == 0
    dx = torch.empty((n, h, w, c1 + c2), dtype=torch.float32, device="cuda")
    assert L.cageo_nn_conv3x3(None, p(dg), cout, 0, None, 0, p(wflip), None, None, p(dx), n,
                              h, w, c1 + c2, 0) == 0
    torch.cuda.synchronize()
    k = (n * h * w) ** 0.5
    assert np.abs(dwt.cpu().numpy() - dw_ref).max() <= 1e-5 * np.abs(dw_ref).max() * k + 1e-6
    # the data gradient of an upsampled input is at full resolution here
    assert np.abs(dx.cpu().numpy() - dx_ref).max() <= 1e-5 * np.abs(dx_ref).max() * cout ** 0.5


FUSED_CASES = [  # n, h, w, c1, up1, c2, cout, pool, head
    (1, 8, 128, 32, 0, 0, 32, True, False),    # 8-row workgroups, 32 channels
    (1, 8, 128, 32, 0, 0, 32, False, True),
    (2, 4, 64, 64, 1, 32, 64, True, False),    # 4-row workgroups, 64 channels
    (1, 4, 64, 32, 0, 0, 64, False, True),
    (1, 12, 64, 32, 0, 0, 32, True, False),    # 4-row workgroups, 32 channels
    (1, 12, 64, 32, 0, 0, 32, False, True),
    (1, 8, 64, 32, 0, 0, 96, True, False),     # two channel blocks: pool only
]


@pytest.mark.parametrize("case", FUSED_CASES)
def test_fused_consumers_equal_their_own_kernels(env, case):
    """cageo_nn_conv3x3_fused: the output, its 2x2 average pool and the 1x1
    head computed in the convolution's launch are the bits of the
    convolution followed by cageo_nn_avgpool2 / cageo_nn_head."""
    torch, sg, L = env
    n, h, w, c1, up1, c2, cout, pool, head = case
    announce("nn fused conv", case)
    r = np.random.default_rng(sum(case))
    h1, w1 = ((h + 1) // 2, (w + 1) // 2) if up1 else (h, w)
    d1 = dev_bits(torch, r.normal(size=(n, h1, w1, c1)))
    d2 = dev_bits(torch, r.normal(size=(n, h, w, c2))) if c2 else None
    dw_ = dev_bits(torch, r.normal(size=(9, cout, c1 + c2)) * 0.1)
    db = torch.from_numpy(r.normal(size=cout).astype(np.float32)).cuda()
    hw = torch.from_numpy(r.normal(size=cout).astype(np.float32)).cuda()
    hb = torch.from_numpy(r.normal(size=1).astype(np.float32)).cuda()
    p = sg._p
    assert L.cageo_nn_conv3x3_fusable(h, w, cout, int(head)) == 1
    assert L.cageo_nn_conv3x3_fusable(h, w + 8, cout, 0) == 0
    ref = torch.empty((n, h, w, cout), dtype=torch.int16, device="cuda")
    assert L.cageo_nn_conv3x3(None, p(d1), c1, up1, p(d2), c2, p(dw_), p(db), p(ref), None, n, h,
                              w, cout, 1) == 0
    if pool:
        ref_pool = torch.empty((n, h // 2, w // 2, cout), dtype=torch.int16, device="cuda")
        assert L.cageo_nn_avgpool2(None, p(ref), p(ref_pool), n, h, w, cout) == 0
        out = torch.full((n, h, w, cout), -1, dtype=torch.int16, device="cuda")
        got = torch.full((n, h // 2, w // 2, cout), -1, dtype=torch.int16, device="cuda")
        assert L.cageo_nn_conv3x3_fused(None, p(d1), c1, up1, p(d2), c2, p(dw_), p(db), p(out),
                                        p(got), None, None, None, 0, n, h, w, cout, 1) == 0
        torch.cuda.synchronize()
        assert torch.equal(out, ref) and torch.equal(got, ref_pool)
    if head:
        ref_prob = torch.empty((n, h, w), dtype=torch.float32, device="cuda")
        assert L.cageo_nn_head(None, p(ref), p(hw), p(hb), p(ref_prob), n * h * w, cout, 1) == 0
        got = torch.full((n, h, w), -1.0, dtype=torch.float32, device="cuda")
        assert L.cageo_nn_conv3x3_fused(None, p(d1), c1, up1, p(d2), c2, p(dw_), p(db), None,
                                        None, p(hw), p(hb), p(got), 1, n, h, w, cout, 1) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, ref_prob)
    else:
        # a head needs the whole channel count in one workgroup (32 or 64)
        expect = 1 if cout in (32, 64) else 0
        assert L.cageo_nn_conv3x3_fusable(h, w, cout, 1) == expect


def test_network_gradients_match_oracle(env):
    """Whole U-Net (base 32) at 32x32 against the float64 oracle with the same
    (bf16-rounded) weights and inputs, twice:
    - with the oracle rounding to bf16 where the kernels store bf16 values
      (activations, pooled tensors, back-propagated gradients).  Fed the same
      inputs, each layer then agrees except for 1-ulp rounding flips in about
      1e-4 of its outputs (FP32 versus FP64 sums near a rounding boundary);
      through 14 layers and the backward pass these flips compound, so the
      per-tensor gradients agree to a relative L2 error <= 0.06 and a cosine
      >= 0.998 (single layers are pinned at 1e-5 by the test above);
    - with the plain float64 oracle: <= 0.15 and cosine >= 0.99, the size of
      the bf16 rounding itself.
    A wrong index or transposition gives errors of order one in both."""
    torch, sg, L = env
    announce("nn network gradients 32x32")
    params = sg.init_params(32, seed=5)
    r = np.random.default_rng(6)
    for k in params:
        if k.endswith(".w"):
            params[k] = sg.bf16_round(params[k])
        else:
            params[k] = r.normal(0, 0.05, params[k].shape).astype(np.float32)
    dem = r.normal(size=(32, 32)).astype(np.float32)
    x = sg.make_inputs(dem, r.random((32, 32)) < 0.1, 1.8, 60.0)[None]
    x[..., :4] = sg.bf16_round(x[..., :4])
    t = r.uniform(0, 1, (1, 32, 32)).astype(np.float32)
    net = sg.UNet(torch, base=32, params=params)
    net.zero_grad()
    net.forward(torch.from_numpy(x).cuda())
    loss = net.backward(torch.from_numpy(t).cuda())
    got = net.grads_numpy()
    for label, rnd, tol_rel, tol_cos in [("bf16-emulating oracle", unet.bf16, 0.06, 0.998),
                                         ("float64 oracle", None, 0.15, 0.99)]:
        ref_loss, ref, _ = unet.forward_backward(params, x, t, rnd=rnd)
        assert abs(loss - ref_loss) <= 0.01 * abs(ref_loss), (label, loss, ref_loss)
        bad = []
        for k in ref:
            a, b = got[k].astype(np.float64).ravel(), ref[k].ravel()
            if k == "enc1a.w":
                a, b = got[k][:, :, :4].ravel().astype(np.float64), ref[k][:, :, :4].ravel()
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            print(f"  {label:22s} {k:10s} rel {rel:.4f} cos {cos:.6f}")
            if not (rel <= tol_rel and cos >= tol_cos):
                bad.append((k, rel, cos))
        assert not bad, (label, bad)


def test_training_reduces_loss(env):
    torch, sg, L = env
    announce("nn training steps 64x64 x4")
    r = np.random.default_rng(7)
    xs, ts = [], []
    for i in range(4):
        dem = np.cumsum(r.normal(size=(64, 64)), axis=1).astype(np.float32)
        xs.append(sg.make_inputs(dem, r.random((64, 64)) < 0.05, 1.5, 80.0))
        # a learnable target: borders where the z-scored DEM is high
        ts.append((xs[-1][..., 0] > 0.5).astype(np.float32))
    x = torch.from_numpy(np.stack(xs)).cuda()
    t = torch.from_numpy(np.stack(ts)).cuda()
    net = sg.UNet(torch, base=32, seed=1)
    losses = []
    for step in range(30):
        net.zero_grad()
        net.forward(x)
        losses.append(net.backward(t))
        net.adam_step(lr=2e-3)
    assert losses[-1] < 0.5 * losses[0], losses
    prob = net.predict(x).cpu().numpy()
    assert prob.shape == (4, 64, 64) and np.all((prob >= 0) & (prob <= 1))


def test_trained_surrogate_gain(env):
    """The packaged weights on a DEM outside the training set (synthetic
    2048^2, seed 777; training used seeds 1000..1063): mean absolute deviation
    from a 1000-realization Monte Carlo <= 0.02 of border probability, the
    check accepts, and the checked prediction takes well under the Monte
    Carlo's time (>= 10x here: the full-size figure is measured by bench.py)."""
    import time
    import cageo_amd as ca
    torch, sg, L = env
    announce("surrogate gain 2048^2 vs 1000-realization MC")
    dem, st = ca.synth_dem(2048, 2048, 777, 10.0, 2000, -9999.0)
    net = sg.load_net(torch)
    sg.checked_prediction(torch, net, dem[:256, :256], st[:256, :256], 1.8, 60.0, 1,
                          check_realizations=2, fallback_realizations=2)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    p_mc = sg.mc_frequency(torch, [dem], [st], [1.8], [60.0], 1000, 11, on_device=True)[0]
    torch.cuda.synchronize()
    mc_s = time.perf_counter() - t0
    q, info = sg.checked_prediction(torch, net, dem, st, 1.8, 60.0, 12)
    sur_s = info["predict_s"] + info["check_s"]
    l1 = float((q - p_mc).abs().mean())
    print(f"  l1 {l1:.5f} mc {mc_s:.3f} s surrogate {sur_s:.3f} s info {info}")
    assert info["accepted"], info
    assert l1 <= 0.02
    assert mc_s >= 10 * sur_s
    assert float(q.min()) >= 0.0 and float(q.max()) <= 1.0


def test_surrogate_fallback(env):
    """A prediction the check rejects is replaced by the full Monte Carlo:
    a network that says 0.993 everywhere (head weights 0, bias 5) is far from
    any check map, so checked_prediction returns exactly the map of
    mc_frequency with the fallback's size and seed (acceptance of a good
    prediction is covered by the gain test)."""
    import cageo_amd as ca
    torch, sg, L = env
    announce("surrogate fallback 256^2")
    dem, st = ca.synth_dem(256, 256, 31, 10.0, 300, -9999.0)
    params = sg.load_params(sg.DEFAULT_WEIGHTS)
    params["head.w"] = np.zeros_like(params["head.w"])
    params["head.b"] = np.array([5.0], np.float32)
    bad = sg.UNet(torch, base=32, params=params)
    m, info = sg.checked_prediction(torch, bad, dem, st, 1.5, 80.0, 40, check_realizations=4,
                                    fallback_realizations=64)
    assert not info["accepted"] and "fallback_s" in info, info
    ref = sg.mc_frequency(torch, [dem], [st], [1.5], [80.0], 64, 41, on_device=True)[0]
    assert torch.equal(m, ref)


def test_native_surrogate_matches_python(env):
    """The C-ABI surrogate (cageo_surrogate_*, C++ composition of the same
    kernels) against the Python composition with the packaged weights:
    bit-identical forward pass on the same bf16 input; predictions from the
    DEM within the rounding of the input planes (mean / deviation in double
    there, float32 here); the Monte Carlo check statistics equal to the
    check statistics (unbiased mean signed error and RMS error, whole map and
    64 x 64 windows) equal to a numpy evaluation of the batched Monte Carlo's
    counts for the same seed."""
    import cageo_amd as ca
    torch, sg, L = env
    announce("native surrogate vs python 256x200")
    dem, st = ca.synth_dem(200, 256, 5, 10.0, 300, -9999.0)
    net = sg.load_net(torch)
    x = sg.device_inputs(torch, dem, st, 1.8, 60.0)  # [1][256][200][32] float32
    xb = net._to_bf16(x)
    ref = net.predict(x)[0]
    out = torch.empty((256, 200), dtype=torch.float32, device="cuda")
    with ca.Surrogate(sg.DEFAULT_WEIGHTS, torch.cuda.current_stream()) as s:
        s.forward(xb, 256, 200, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        ddem, dst = torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda()
        p = torch.empty((256, 200), dtype=torch.float32, device="cuda")
        s.predict(ddem, dst, -9999.0, -9999.0, 1.8, 60.0, p)
        torch.cuda.synchronize()
        assert float((p - ref).abs().max()) <= 0.02 and float((p - ref).abs().mean()) <= 1e-3
        r = ca.filter_radius(60.0, 10.0)
        with ca.Context(200, 256, torch.cuda.current_stream()) as ctx:
            ctx.set_filter(ca.GAUSSIAN, r, ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r))
            ctx.mc_setup(ddem, dst, -9999.0, -9999.0, 2, 0.0, 1.8)
            scratch = torch.empty((256, 200), dtype=torch.float32, device="cuda")
            acc, bias, rms = s.check(ctx, p, 4, 21, 0.01, 0.05, scratch)
            wacc, wbias, wrms = ctx.check_probability_map(p, 4, 21, 0.01, 0.05, scratch,
                                                          window=64)
        # the same statistics from the batched Monte Carlo's 4 realizations
        chk = sg.mc_frequency(torch, [dem], [st], [1.8], [60.0], 4, 21, on_device=True)[0]
        f, q = chk.double().cpu().numpy(), p.double().cpu().numpy()
        u = (f - q) ** 2 - f * (1 - f) / 3.0
        b_py, r_py = (f - q).mean(), np.sqrt(max(u.mean(), 0.0))
        assert abs(bias - b_py) <= 1e-9 and abs(rms - r_py) <= 1e-9, (bias, b_py, rms, r_py)
        assert acc == (abs(b_py) <= 0.01 and r_py <= 0.05)
        assert wacc.shape == (4, 4)  # 256 / 64 rows, ceil(200 / 64) columns
        for wy in range(4):
            for wx in range(4):
                sl = (slice(64 * wy, 64 * wy + 64), slice(64 * wx, 64 * wx + 64))
                bw = (f - q)[sl].mean()
                rw = np.sqrt(max(u[sl].mean(), 0.0))
                assert abs(wbias[wy, wx] - bw) <= 1e-9 and abs(wrms[wy, wx] - rw) <= 1e-9
                assert wacc[wy, wx] == (abs(bw) <= 0.01 and rw <= 0.05)


def test_refinement_replaces_exactly_the_failing_windows(env):
    """cageo_refine_probability_map: windows flagged as failing take the
    frequency count / n of an n-realization Monte Carlo of the whole DEM
    (the same counts cageo_run_realizations gives for that seed), ragged
    edge windows included; accepted windows keep the prediction bit for
    bit."""
    import cageo_amd as ca
    torch, sg, L = env
    announce("refinement of flagged windows 200x150, window 64")
    dem, st = ca.synth_dem(200, 150, 17, 10.0, 300, -9999.0)
    ddem, dst = torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda()
    r = ca.filter_radius(60.0, 10.0)
    with ca.Context(200, 150, torch.cuda.current_stream()) as ctx:
        ctx.set_filter(ca.GAUSSIAN, r, ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r))
        ctx.mc_setup(ddem, dst, -9999.0, -9999.0, 2, 0.0, 1.8)
        nx, ny = ctx.check_windows(64)
        assert (nx, ny) == (4, 3)  # ceil(200 / 64), ceil(150 / 64)
        q0 = torch.rand((150, 200), device="cuda", generator=None)
        q = q0.clone()
        acc = np.ones((ny, nx), bool)
        acc[0, 1] = acc[2, 3] = acc[1, 0] = False  # an edge window, the corner, an inner one
        scratch = torch.empty((150, 200), dtype=torch.float32, device="cuda")
        ctx.refine_probability_map(q, acc, 12, 33, scratch, window=64)
        counts = torch.zeros((150, 200), dtype=torch.float32, device="cuda")
        ctx.run_realizations(counts, 12, 33, 0)
        torch.cuda.synchronize()
        # the library scales by the float reciprocal of n
        freq = (counts.cpu().numpy() * (np.float32(1) / np.float32(12))).astype(np.float32)
        got, before = q.cpu().numpy(), q0.cpu().numpy()
    for wy in range(ny):
        for wx in range(nx):
            sl = (slice(64 * wy, 64 * wy + 64), slice(64 * wx, 64 * wx + 64))
            want = before[sl] if acc[wy, wx] else freq[sl]
            assert np.array_equal(got[sl], want), (wy, wx)
