"""Border-probability surrogate: a 4-level U-Net on libcageo_hip's BF16 MFMA kernels.

No reference counterpart: the reference only runs the Monte Carlo.  The
network maps four input planes -- the z-scored DEM, the stream mask, the
error RMSE and the practical range (as constant planes) -- to the per-cell
probability that a realization puts a catchment border there, i.e. the
Monte Carlo's border count divided by its number of realizations.  The
Monte Carlo produces its training maps and stays both the judge of its
output (the error check) and the fallback.

Architecture (base b = 32 channels, 3x3 convolutions, ReLU):
  enc1: 4(->32 padded) -> b -> b          at H
  enc2: avgpool, b -> 2b -> 2b            at H/2
  enc3: avgpool, 2b -> 4b -> 4b           at H/4
  mid : avgpool, 4b -> 8b -> 8b           at H/8
  dec3: [up(mid) | enc3] -> 4b -> 4b      at H/4
  dec2: [up(dec3) | enc2] -> 2b -> 2b     at H/2
  dec1: [up(dec2) | enc1] -> b -> b       at H
  head: 1x1, b -> 1, sigmoid
Weights and activations are BF16, accumulation FP32 (the MFMA kernels);
master weights, gradients and Adam moments are FP32.  Every layer runs on
the library's kernels; torch only holds device memory and the stream.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import lib, _check

IN_CHANNELS = 4
IN_PAD = 32  # the first convolution reads 32 channels (4 used)

# Weights trained by `python -m cageo_amd.train_surrogate` (320 synthetic DEMs
# of 4096^2 cut into 20480 tiles of 512^2, 200-realization Monte Carlo
# labels, Gaussian filter, 10000 Adam steps of 8 tiles).
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(__file__), "weights", "surrogate_b32.bin")


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


_SIGS = None


def _nn():
    global _SIGS
    L = lib()
    if _SIGS is None:
        vp, i, f, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
        sigs = {
            "cageo_nn_conv3x3": [vp, vp, i, i, vp, i, vp, vp, vp, vp, i, i, i, i, i],
            "cageo_nn_conv3x3_fused": [vp, vp, i, i, vp, i, vp, vp, vp, vp, vp, vp, vp, i,
                                       i, i, i, i, i],
            "cageo_nn_conv3x3_fusable": [i, i, i, i],
            "cageo_nn_conv3x3_wgrad": [vp, vp, i, i, vp, i, vp, vp, i, i, i, i],
            "cageo_nn_flip_transpose": [vp, vp, vp, i, i],
            "cageo_nn_avgpool2": [vp, vp, vp, i, i, i, i],
            "cageo_nn_avgpool2_back": [vp, vp, vp, i, i, i, i],
            "cageo_nn_upsample2_back": [vp, vp, vp, i, i, i, i, i, i],
            "cageo_nn_slice_add": [vp, vp, vp, ll, i, i, i, i],
            "cageo_nn_relu_back": [vp, vp, vp, vp, ll],
            "cageo_nn_bias_grad": [vp, vp, vp, ll, i],
            "cageo_nn_head": [vp, vp, vp, vp, vp, ll, i, i],
            "cageo_nn_head_back": [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, ll, i],
            "cageo_nn_adam": [vp, vp, vp, vp, vp, vp, ll, f, f, f, f, i],
            "cageo_nn_to_bf16": [vp, vp, vp, ll],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _SIGS = sigs
    return L


def layer_spec(base=32):
    """Convolution layers in order: (name, in1 channels, upsampled, in2 channels, out)."""
    b = base
    return [
        ("enc1a", IN_PAD, 0, 0, b), ("enc1b", b, 0, 0, b),
        ("enc2a", b, 0, 0, 2 * b), ("enc2b", 2 * b, 0, 0, 2 * b),
        ("enc3a", 2 * b, 0, 0, 4 * b), ("enc3b", 4 * b, 0, 0, 4 * b),
        ("mida", 4 * b, 0, 0, 8 * b), ("midb", 8 * b, 0, 0, 8 * b),
        ("dec3a", 8 * b, 1, 4 * b, 4 * b), ("dec3b", 4 * b, 0, 0, 4 * b),
        ("dec2a", 4 * b, 1, 2 * b, 2 * b), ("dec2b", 2 * b, 0, 0, 2 * b),
        ("dec1a", 2 * b, 1, b, b), ("dec1b", b, 0, 0, b),
    ]


# Resolution level (0 = full, k = 1/2^k) of each convolution layer.
LAYER_LEVEL = {"enc1a": 0, "enc1b": 0, "enc2a": 1, "enc2b": 1, "enc3a": 2, "enc3b": 2,
               "mida": 3, "midb": 3, "dec3a": 2, "dec3b": 2, "dec2a": 1, "dec2b": 1,
               "dec1a": 0, "dec1b": 0}


def forward_flops(h, w, base=32, in_channels=IN_PAD):
    """Multiply-add FLOPs (2 per MAC) of one forward pass over an h x w input,
    the first layer over `in_channels` input channels (IN_PAD as the padded
    layer executes it, IN_CHANNELS for the products that count)."""
    total = 0
    for name, c1, _, c2, co in layer_spec(base):
        if name == "enc1a":
            c1 = in_channels
        lv = LAYER_LEVEL[name]
        total += 2 * 9 * (c1 + c2) * co * (h >> lv) * (w >> lv)
    return total + 2 * base * h * w  # the 1x1 head


def init_params(base=32, seed=0):
    """He-normal weights ([9][out][in], the padded input channels zero) and
    zero biases, as float32 numpy arrays keyed by layer name."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, c1, _, c2, co in layer_spec(base):
        cin = c1 + c2
        fan_in = 9 * (IN_CHANNELS if name == "enc1a" else cin)
        w = rng.normal(0.0, math.sqrt(2.0 / fan_in), (9, co, cin)).astype(np.float32)
        if name == "enc1a":
            w[:, :, IN_CHANNELS:] = 0.0
        params[name + ".w"] = w
        params[name + ".b"] = np.zeros(co, np.float32)
    params["head.w"] = rng.normal(0.0, math.sqrt(1.0 / base), base).astype(np.float32)
    params["head.b"] = np.zeros(1, np.float32)
    return params


def bf16_round(a):
    """Round float32 to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def make_inputs(dem, stream_mask, rmse, practical_range, nodata=-9999.0, channels=IN_PAD):
    """The 4 input planes, NHWC float32 [H][W][channels] (channels >= 4, the
    rest zero): z-scored DEM (nodata cells 0), stream mask (0/1), RMSE / 3 m,
    practical range / 200 m."""
    dem = np.asarray(dem, np.float32)
    valid = dem != nodata
    mu = float(dem[valid].mean()) if valid.any() else 0.0
    sd = float(dem[valid].std()) if valid.any() else 1.0
    x = np.zeros(dem.shape + (channels,), np.float32)
    x[..., 0] = np.where(valid, (dem - mu) / max(sd, 1e-6), 0.0)
    x[..., 1] = np.asarray(stream_mask, np.float32)
    x[..., 2] = rmse / 3.0
    x[..., 3] = practical_range / 200.0
    return x


class UNet:
    """The surrogate network on one device; `torch` supplies memory/stream."""

    def __init__(self, torch, base=32, seed=0, params=None):
        self.torch = torch
        self.base = base
        self.spec = layer_spec(base)
        self.dev = torch.device("cuda")
        self.stream = torch.cuda.current_stream().cuda_stream
        host = params if params is not None else init_params(base, seed)
        T = torch
        self.p32, self.pbf, self.grad, self.m, self.v = {}, {}, {}, {}, {}
        for k, a in host.items():
            t = T.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.dev)
            self.p32[k] = t
            self.grad[k] = T.zeros_like(t)
            self.m[k] = T.zeros_like(t)
            self.v[k] = T.zeros_like(t)
            if k.endswith(".w") and not k.startswith("head"):
                self.pbf[k] = T.empty(t.shape, dtype=T.int16, device=self.dev)
        self.step_count = 0
        self.refresh_bf16()

    # -- parameters -----------------------------------------------------------
    def refresh_bf16(self):
        L = _nn()
        for k, t in self.pbf.items():
            _check(L.cageo_nn_to_bf16(self.stream, _p(self.p32[k]), _p(t), t.numel()), "to_bf16")

    def params_numpy(self):
        return {k: t.detach().cpu().numpy().copy() for k, t in self.p32.items()}

    def grads_numpy(self):
        return {k: t.detach().cpu().numpy().copy() for k, t in self.grad.items()}

    def zero_grad(self):
        for t in self.grad.values():
            t.zero_()

    # -- forward ----------------------------------------------------------------
    def _to_bf16(self, x32):
        T = self.torch
        out = T.empty(x32.shape, dtype=T.int16, device=self.dev)
        _check(_nn().cageo_nn_to_bf16(self.stream, _p(x32), _p(out), x32.numel()), "to_bf16")
        return out

    def _conv(self, name, in1, c1, up1, in2, c2, n, h, w, relu=True, keep_f32=False):
        T = self.torch
        co = self.p32[name + ".b"].numel()
        out = T.empty((n, h, w, co), dtype=T.int16, device=self.dev)
        _check(_nn().cageo_nn_conv3x3(self.stream, _p(in1), c1, up1, _p(in2), c2,
                                      _p(self.pbf[name + ".w"]), _p(self.p32[name + ".b"]),
                                      _p(out), None, n, h, w, co, int(relu)), name)
        return out

    def _pool(self, x, n, h, w, c):
        T = self.torch
        out = T.empty((n, h // 2, w // 2, c), dtype=T.int16, device=self.dev)
        _check(_nn().cageo_nn_avgpool2(self.stream, _p(x), _p(out), n, h, w, c), "avgpool")
        return out

    def forward(self, x32, keep=True, sigmoid=False):
        """x32: float32 [N][H][W][IN_PAD] on the device (H, W divisible by 8).
        Returns logits [N][H][W] (float32), or probabilities with sigmoid."""
        T = self.torch
        n, h, w, _ = x32.shape
        b = self.base
        acts = {"x": self._to_bf16(x32)}
        e1a = self._conv("enc1a", acts["x"], IN_PAD, 0, None, 0, n, h, w)
        e1 = self._conv("enc1b", e1a, b, 0, None, 0, n, h, w)
        p1 = self._pool(e1, n, h, w, b)
        e2a = self._conv("enc2a", p1, b, 0, None, 0, n, h // 2, w // 2)
        e2 = self._conv("enc2b", e2a, 2 * b, 0, None, 0, n, h // 2, w // 2)
        p2 = self._pool(e2, n, h // 2, w // 2, 2 * b)
        e3a = self._conv("enc3a", p2, 2 * b, 0, None, 0, n, h // 4, w // 4)
        e3 = self._conv("enc3b", e3a, 4 * b, 0, None, 0, n, h // 4, w // 4)
        p3 = self._pool(e3, n, h // 4, w // 4, 4 * b)
        ma = self._conv("mida", p3, 4 * b, 0, None, 0, n, h // 8, w // 8)
        mid = self._conv("midb", ma, 8 * b, 0, None, 0, n, h // 8, w // 8)
        d3a = self._conv("dec3a", mid, 8 * b, 1, e3, 4 * b, n, h // 4, w // 4)
        d3 = self._conv("dec3b", d3a, 4 * b, 0, None, 0, n, h // 4, w // 4)
        d2a = self._conv("dec2a", d3, 4 * b, 1, e2, 2 * b, n, h // 2, w // 2)
        d2 = self._conv("dec2b", d2a, 2 * b, 0, None, 0, n, h // 2, w // 2)
        d1a = self._conv("dec1a", d2, 2 * b, 1, e1, b, n, h, w)
        d1 = self._conv("dec1b", d1a, b, 0, None, 0, n, h, w)
        logit = T.empty((n, h, w), dtype=T.float32, device=self.dev)
        _check(_nn().cageo_nn_head(self.stream, _p(d1), _p(self.p32["head.w"]),
                                   _p(self.p32["head.b"]), _p(logit), n * h * w, b,
                                   int(sigmoid)), "head")
        if keep:
            self._acts = dict(x=acts["x"], e1a=e1a, e1=e1, p1=p1, e2a=e2a, e2=e2, p2=p2, e3a=e3a,
                              e3=e3, p3=p3, ma=ma, mid=mid, d3a=d3a, d3=d3, d2a=d2a, d2=d2,
                              d1a=d1a, d1=d1, logit=logit, shape=(n, h, w))
        return logit

    def predict(self, x32):
        """Probability map [N][H][W] (float32, on the device)."""
        return self.forward(x32, keep=False, sigmoid=True)

    # -- backward ---------------------------------------------------------------
    def _conv_back(self, name, g_out_f32, act_out, in1, c1, up1, in2, c2, n, h, w, need_dx=True):
        """Gradients of a ReLU conv layer from d(output) (fp32).  Adds dW, db;
        returns d(input) (fp32 [n][h][w][c1+c2], at the layer's resolution)."""
        T, L = self.torch, _nn()
        co = self.p32[name + ".b"].numel()
        g = T.empty((n, h, w, co), dtype=T.int16, device=self.dev)
        _check(L.cageo_nn_relu_back(self.stream, _p(g_out_f32), _p(act_out), _p(g), g.numel()),
               "relu_back")
        _check(L.cageo_nn_bias_grad(self.stream, _p(g), _p(self.grad[name + ".b"]), n * h * w, co),
               "bias_grad")
        _check(L.cageo_nn_conv3x3_wgrad(self.stream, _p(in1), c1, up1, _p(in2), c2, _p(g),
                                        _p(self.grad[name + ".w"]), n, h, w, co), "wgrad")
        if not need_dx:
            return None
        cin = c1 + c2
        wt = T.empty((9, cin, co), dtype=T.int16, device=self.dev)
        _check(L.cageo_nn_flip_transpose(self.stream, _p(self.pbf[name + ".w"]), _p(wt), co, cin),
               "flip")
        dx = T.empty((n, h, w, cin), dtype=T.float32, device=self.dev)
        _check(L.cageo_nn_conv3x3(self.stream, _p(g), co, 0, None, 0, _p(wt), None, None, _p(dx),
                                  n, h, w, cin, 0), name + " dgrad")
        return dx

    def backward(self, target):
        """BCE loss against `target` (float32 [N][H][W] device tensor of
        frequencies); accumulates parameter gradients; returns the loss."""
        T, L = self.torch, _nn()
        a = self._acts
        n, h, w = a["shape"]
        b = self.base
        P = n * h * w
        q = T.empty((n, h, w), dtype=T.float32, device=self.dev)
        dd1 = T.empty((n, h, w, b), dtype=T.float32, device=self.dev)
        loss = T.zeros(1, dtype=T.float32, device=self.dev)
        _check(L.cageo_nn_head_back(self.stream, _p(a["d1"]), _p(self.p32["head.w"]),
                                    _p(a["logit"]), _p(target), _p(q), _p(dd1),
                                    _p(self.grad["head.w"]), _p(self.grad["head.b"]), _p(loss),
                                    P, b), "head_back")
        h2, w2, h4, w4, h8, w8 = h // 2, w // 2, h // 4, w // 4, h // 8, w // 8

        def zeros(*s):
            return T.zeros(s, dtype=T.float32, device=self.dev)

        dd1a = self._conv_back("dec1b", dd1, a["d1"], a["d1a"], b, 0, None, 0, n, h, w)
        dcat1 = self._conv_back("dec1a", dd1a, a["d1a"], a["d2"], 2 * b, 1, a["e1"], b, n, h, w)
        de1 = zeros(n, h, w, b)
        _check(L.cageo_nn_slice_add(self.stream, _p(dcat1), _p(de1), P, b, 3 * b, 2 * b, 1), "split")
        dd2 = zeros(n, h2, w2, 2 * b)
        _check(L.cageo_nn_upsample2_back(self.stream, _p(dcat1), _p(dd2), n, h, w, 2 * b, 3 * b, 0),
               "up_back")
        dd2a = self._conv_back("dec2b", dd2, a["d2"], a["d2a"], 2 * b, 0, None, 0, n, h2, w2)
        dcat2 = self._conv_back("dec2a", dd2a, a["d2a"], a["d3"], 4 * b, 1, a["e2"], 2 * b, n, h2, w2)
        de2 = zeros(n, h2, w2, 2 * b)
        _check(L.cageo_nn_slice_add(self.stream, _p(dcat2), _p(de2), n * h2 * w2, 2 * b, 6 * b,
                                    4 * b, 1), "split")
        dd3 = zeros(n, h4, w4, 4 * b)
        _check(L.cageo_nn_upsample2_back(self.stream, _p(dcat2), _p(dd3), n, h2, w2, 4 * b, 6 * b,
                                         0), "up_back")
        dd3a = self._conv_back("dec3b", dd3, a["d3"], a["d3a"], 4 * b, 0, None, 0, n, h4, w4)
        dcat3 = self._conv_back("dec3a", dd3a, a["d3a"], a["mid"], 8 * b, 1, a["e3"], 4 * b, n, h4,
                                w4)
        de3 = zeros(n, h4, w4, 4 * b)
        _check(L.cageo_nn_slice_add(self.stream, _p(dcat3), _p(de3), n * h4 * w4, 4 * b, 12 * b,
                                    8 * b, 1), "split")
        dmid = zeros(n, h8, w8, 8 * b)
        _check(L.cageo_nn_upsample2_back(self.stream, _p(dcat3), _p(dmid), n, h4, w4, 8 * b, 12 * b,
                                         0), "up_back")
        dma = self._conv_back("midb", dmid, a["mid"], a["ma"], 8 * b, 0, None, 0, n, h8, w8)
        dp3 = self._conv_back("mida", dma, a["ma"], a["p3"], 4 * b, 0, None, 0, n, h8, w8)
        _check(L.cageo_nn_avgpool2_back(self.stream, _p(dp3), _p(de3), n, h4, w4, 4 * b), "pool_back")
        de3a = self._conv_back("enc3b", de3, a["e3"], a["e3a"], 4 * b, 0, None, 0, n, h4, w4)
        dp2 = self._conv_back("enc3a", de3a, a["e3a"], a["p2"], 2 * b, 0, None, 0, n, h4, w4)
        _check(L.cageo_nn_avgpool2_back(self.stream, _p(dp2), _p(de2), n, h2, w2, 2 * b), "pool_back")
        de2a = self._conv_back("enc2b", de2, a["e2"], a["e2a"], 2 * b, 0, None, 0, n, h2, w2)
        dp1 = self._conv_back("enc2a", de2a, a["e2a"], a["p1"], b, 0, None, 0, n, h2, w2)
        _check(L.cageo_nn_avgpool2_back(self.stream, _p(dp1), _p(de1), n, h, w, b), "pool_back")
        de1a = self._conv_back("enc1b", de1, a["e1"], a["e1a"], b, 0, None, 0, n, h, w)
        self._conv_back("enc1a", de1a, a["e1a"], a["x"], IN_PAD, 0, None, 0, n, h, w, need_dx=False)
        # the padded input channels of the first layer are not trained
        self.grad["enc1a.w"][:, :, IN_CHANNELS:] = 0.0
        return float(loss.item())

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _nn()
        self.step_count += 1
        for k, p in self.p32.items():
            _check(L.cageo_nn_adam(self.stream, _p(p), _p(self.grad[k]), _p(self.m[k]),
                                   _p(self.v[k]), _p(self.pbf.get(k)), p.numel(), lr, b1, b2, eps,
                                   self.step_count), "adam")


# ---- Monte Carlo labels, training, checked prediction -------------------------

def mc_frequency(torch, dems, streams, rmses, ranges, realizations, seed, nodata=-9999.0,
                 on_device=False, model=None):
    """Border frequency (count / realizations) of each DEM by the batched Monte
    Carlo (error field with the given RMSE and practical range, Gaussian
    filter unless `model`), as float32 numpy arrays (device tensors with
    on_device)."""
    from . import Batch, BATCH_GRAPH, GAUSSIAN
    model = GAUSSIAN if model is None else model
    shapes = [d.shape for d in dems]
    keep = []
    with Batch(shapes) as b:
        for t, (d, s, e, pr) in enumerate(zip(dems, streams, rmses, ranges)):
            dd, ss = torch.from_numpy(d).cuda(), torch.from_numpy(s).cuda()
            keep.append((dd, ss))
            b.setup_tile(t, dd, ss, nodata, nodata, 2, model, float(pr), 10.0, 0.0, float(e),
                         seed + t)
        mcs = [torch.zeros(sh, dtype=torch.float32, device="cuda") for sh in shapes]
        b.run(mcs, realizations, 0, BATCH_GRAPH)
        if on_device:
            return [m / float(realizations) for m in mcs]
        out = [m.cpu().numpy() / np.float32(realizations) for m in mcs]
    return out


def training_set(torch, n_dems=16, size=2048, crop=512, realizations=200, seed=1000,
                 acc_threshold=2000, nodata=-9999.0, model=None, rmse_range=(0.5, 3.0),
                 pr_range=(30.0, 200.0), log=None, log_every=8, fast_convolution=False):
    """Inputs [N][crop][crop][4] and labels [N][crop][crop] as float32 device
    tensors, N = n_dems * (size // crop)^2: synthetic DEMs of the evaluation
    generator (seeds seed + k), each with its own RMSE ~ U(rmse_range) m and
    practical range ~ U(pr_range) m, labelled by `realizations` Monte Carlo
    realizations of the whole DEM on one context (Gaussian filter unless
    `model`), input planes built on the device (the DEM z-scored over the
    whole raster, as cageo_surrogate_predict does), then cut into crops.
    With fast_convolution the labels' error fields come from the fast
    convolution mode (within 1e-5 of the exact field).  The host generates
    the next DEM while the device labels the current one."""
    from concurrent.futures import ThreadPoolExecutor
    from . import CONV_FAST, Context, GAUSSIAN, create_filter, filter_radius, synth_dem
    model = GAUSSIAN if model is None else model
    rng = np.random.default_rng(seed)
    rmses = [float(rng.uniform(*rmse_range)) for _ in range(n_dems)]
    ranges = [float(rng.uniform(*pr_range)) for _ in range(n_dems)]
    g = size // crop
    per = g * g
    T = torch
    X = T.empty((n_dems * per, crop, crop, IN_CHANNELS), dtype=T.float32, device="cuda")
    Y = T.empty((n_dems * per, crop, crop), dtype=T.float32, device="cuda")
    mc = T.zeros((size, size), dtype=T.float32, device="cuda")
    gen = lambda k: synth_dem(size, size, seed + k, 10.0, acc_threshold, nodata)
    with ThreadPoolExecutor(1) as pool, Context(size, size, T.cuda.current_stream()) as ctx:
        nxt = pool.submit(gen, 0)
        for k in range(n_dems):
            d, st = nxt.result()
            if k + 1 < n_dems:
                nxt = pool.submit(gen, k + 1)
            dd, ss = T.from_numpy(d).cuda(), T.from_numpy(st).cuda()
            r = filter_radius(ranges[k], 10.0)
            ctx.set_filter(model, r, create_filter(model, ranges[k], 10.0, r))
            if fast_convolution:
                ctx.set_conv_mode(CONV_FAST)
            ctx.mc_setup(dd, ss, nodata, nodata, 2, 0.0, rmses[k])
            mc.zero_()
            ctx.run_realizations(mc, realizations, seed + k, 0)
            valid = dd != nodata
            v = dd[valid].double()
            mu = float(v.mean()) if v.numel() else 0.0
            sd = max(float(v.std(unbiased=False)) if v.numel() else 1.0, 1e-6)
            planes = T.empty((size, size, IN_CHANNELS), dtype=T.float32, device="cuda")
            planes[..., 0] = T.where(valid, (dd - mu) / sd, T.zeros_like(dd))
            planes[..., 1] = (ss != nodata).float()
            planes[..., 2] = rmses[k] / 3.0
            planes[..., 3] = ranges[k] / 200.0
            X[k * per:(k + 1) * per] = (planes.view(g, crop, g, crop, IN_CHANNELS)
                                        .permute(0, 2, 1, 3, 4).reshape(per, crop, crop, IN_CHANNELS))
            Y[k * per:(k + 1) * per] = ((mc / float(realizations)).view(g, crop, g, crop)
                                        .permute(0, 2, 1, 3).reshape(per, crop, crop))
            if log and ((k + 1) % log_every == 0 or k + 1 == n_dems):
                log(f"labels: {k + 1}/{n_dems} DEMs ({(k + 1) * per} tiles)")
    return X, Y, dict(rmse=rmses, practical_range=ranges)


def train(torch, x4, y, steps=1500, batch=8, lr=2e-3, seed=0, base=32, log_every=500,
          log=print, lr_floor=0.02):
    """Adam on the BCE of the surrogate against Monte Carlo frequencies
    through the C ABI (cageo_surrogate_train): batches of `batch` tiles with
    random dihedral transforms, linear warm-up then cosine decay from lr.
    x4 [N][h][w][4] and y [N][h][w] (device tensors or numpy).  Runs in
    pieces of `log_every` steps with a progress line each.  Returns the
    Trainer and the loss history."""
    from . import Trainer
    T = torch
    if isinstance(x4, np.ndarray):
        x4 = T.from_numpy(np.ascontiguousarray(x4, np.float32)).cuda()
    if isinstance(y, np.ndarray):
        y = T.from_numpy(np.ascontiguousarray(y, np.float32)).cuda()
    tr = Trainer(base, seed=seed, stream=T.cuda.current_stream())
    hist = []
    for first in range(0, steps, log_every):
        n = min(log_every, steps - first)
        h = tr.train(x4.contiguous(), y.contiguous(), batch, lr, lr_floor, first, n, steps, seed)
        hist.extend(float(v) for v in h)
        if log:
            log(f"step {first + n} loss {np.mean(h[-min(n, 100):]):.5f}")
    return tr, hist


WEIGHTS_MAGIC = b"CGSURR01"


def save_params(net_or_params, path):
    """Weights -> file.  `.bin`: the library's format (read by
    cageo_surrogate_create): magic "CGSURR01", int32 base, int32 count, then
    per tensor int32 name length, name, int32 dtype (0 float32, 1 bfloat16
    bit patterns), int32 ndim, int32 dims, little-endian data.  `.npz`:
    the same tensors for numpy.  Convolution weights are stored as bfloat16
    (the kernels read bf16, so inference is unchanged), biases and the head
    as float32."""
    import struct
    params = net_or_params.params_numpy() if hasattr(net_or_params, "params_numpy") else net_or_params
    base = int(params["enc1b.b"].size)
    names = [n + sfx for n, *_ in layer_spec(base) for sfx in (".w", ".b")] + ["head.w", "head.b"]
    if str(path).endswith(".bin"):
        with open(path, "wb") as f:
            f.write(WEIGHTS_MAGIC)
            f.write(struct.pack("<ii", base, len(names)))
            for k in names:
                a = np.asarray(params[k], np.float32)
                bf = k.endswith(".w") and not k.startswith("head")
                data = ((bf16_round(a).view(np.uint32) >> 16).astype("<u2") if bf
                        else a.astype("<f4"))
                nb = k.encode()
                f.write(struct.pack("<i", len(nb)) + nb)
                f.write(struct.pack("<ii", 1 if bf else 0, a.ndim))
                f.write(struct.pack("<%di" % a.ndim, *a.shape))
                f.write(data.tobytes())
        return
    out = {}
    for k in names:
        v = params[k]
        if k.endswith(".w") and not k.startswith("head"):
            out[k.replace(".", "__") + "__bf16"] = (bf16_round(v).view(np.uint32) >> 16).astype(np.uint16)
        else:
            out[k.replace(".", "__")] = v
    np.savez_compressed(path, **out)


def load_params(path):
    """Weights from a `.bin` (library format) or `.npz` file, as float32."""
    import struct
    params = {}
    if str(path).endswith(".bin"):
        with open(path, "rb") as f:
            if f.read(8) != WEIGHTS_MAGIC:
                raise ValueError(f"{path}: not a surrogate weights file")
            _, count = struct.unpack("<ii", f.read(8))
            for _ in range(count):
                (ln,) = struct.unpack("<i", f.read(4))
                name = f.read(ln).decode()
                dtype, ndim = struct.unpack("<ii", f.read(8))
                shape = struct.unpack("<%di" % ndim, f.read(4 * ndim))
                nel = int(np.prod(shape))
                if dtype == 1:
                    raw = np.frombuffer(f.read(2 * nel), "<u2").astype(np.uint32) << 16
                    params[name] = raw.view(np.float32).reshape(shape)
                else:
                    params[name] = np.frombuffer(f.read(4 * nel), "<f4").reshape(shape).copy()
        return params
    z = np.load(path)
    for k in z.files:
        if k.endswith("__bf16"):
            params[k[:-6].replace("__", ".")] = (z[k].astype(np.uint32) << 16).view(np.float32)
        else:
            params[k.replace("__", ".")] = z[k]
    return params


def load_net(torch, path=None):
    """The trained surrogate (DEFAULT_WEIGHTS unless `path`) on the device."""
    params = load_params(path or DEFAULT_WEIGHTS)
    return UNet(torch, base=int(params["enc1b.b"].size), params=params)


def expected_l1_of_sample(q, k):
    """Mean over cells of E|B/k - q| with B ~ Binomial(k, q): the L1 distance a
    k-realization Monte Carlo map has from an exact probability map q
    (numpy array or device tensor)."""
    from math import comb
    e = 0.0
    for j in range(k + 1):
        e = e + comb(k, j) * q ** j * (1 - q) ** (k - j) * abs(j / k - q)
    return float(e.mean())


def device_inputs(torch, dem, stream, rmse, practical_range, nodata=-9999.0):
    """make_inputs on the device: [1][H][W][IN_PAD] float32 from host rasters."""
    T = torch
    d = T.from_numpy(np.ascontiguousarray(dem, np.float32)).cuda()
    s = T.from_numpy(np.ascontiguousarray(stream, np.float32)).cuda()
    valid = d != nodata
    dv = d[valid]
    mu = dv.mean() if dv.numel() else T.zeros((), device="cuda")
    sd = dv.std(unbiased=False).clamp_min(1e-6) if dv.numel() else T.ones((), device="cuda")
    x = T.zeros((1,) + tuple(d.shape) + (IN_PAD,), dtype=T.float32, device="cuda")
    x[0, ..., 0] = T.where(valid, (d - mu) / sd, T.zeros_like(d))
    x[0, ..., 1] = (s != nodata).float()
    x[0, ..., 2] = rmse / 3.0
    x[0, ..., 3] = practical_range / 200.0
    return x


# RMS / L1 ratio of the packaged network's per-cell errors against
# 1000-realization Monte Carlo maps (RMS net of the maps' own sampling
# variance), on synthetic 4096^2 DEMs outside its training set (seeds
# 20240601, 20240611, 20240621, Gaussian PR 60 m, RMSE 1.8 m: 2.44 to 2.46;
# L1 0.0158 to 0.0161).
ERROR_RMS_L1_RATIO = 2.46


def check_tolerances(l1_target=0.02, rms_l1_ratio=ERROR_RMS_L1_RATIO, margin=1.2):
    """(bias_tol, rms_tol) of the Monte Carlo check for an L1 target: the mean
    signed error at most half the target, and the RMS error at most `margin`
    times the RMS of a map at the target L1 whose errors are shaped like the
    network's (rms_l1_ratio x target).  The check estimates both without bias
    from a few realizations (Context.check_probability_map)."""
    return 0.5 * l1_target, margin * rms_l1_ratio * l1_target


def checked_prediction(torch, net, dem, stream, rmse, practical_range, seed,
                       check_realizations=8, fallback_realizations=1000, l1_target=0.02,
                       nodata=-9999.0, model=None, prediction=None, ctx=None):
    """Surrogate map with its Monte Carlo check.  `check_realizations`
    realizations of the Monte Carlo give unbiased estimates of the map's mean
    signed error and RMS error (Context.check_probability_map); when either
    exceeds check_tolerances(l1_target), the prediction is rejected and the
    `fallback_realizations` Monte Carlo (seed + 1) is returned instead.
    `prediction` replaces the network's map (tests of the check itself);
    `ctx` is a Context already set up for this DEM and error model.
    Returns (device map, info)."""
    import time
    from . import Context, GAUSSIAN, create_filter, filter_radius
    T = torch
    model = GAUSSIAN if model is None else model
    T.cuda.synchronize()
    t0 = time.perf_counter()
    h, w = dem.shape
    if prediction is None:
        x = device_inputs(torch, dem, stream, rmse, practical_range, nodata)
        ph, pw = -h % 8, -w % 8  # the network halves three times
        if ph or pw:
            xp = T.zeros((1, h + ph, w + pw, IN_PAD), dtype=T.float32, device="cuda")
            xp[:, :h, :w] = x
            x = xp
        q = net.predict(x)[0, :h, :w].contiguous()
    else:
        q = prediction
    T.cuda.synchronize()
    t1 = time.perf_counter()
    own = ctx is None
    if own:
        r = filter_radius(practical_range, 10.0)
        ctx = Context(w, h, T.cuda.current_stream())
        ctx.set_filter(model, r, create_filter(model, practical_range, 10.0, r))
        ctx.mc_setup(T.from_numpy(np.ascontiguousarray(dem, np.float32)).cuda(),
                     T.from_numpy(np.ascontiguousarray(stream, np.float32)).cuda(), nodata,
                     nodata, 2, 0.0, rmse)
    try:
        scratch = T.empty((h, w), dtype=T.float32, device="cuda")
        bias_tol, rms_tol = check_tolerances(l1_target)
        acc, b, rms = ctx.check_probability_map(q, check_realizations, seed, bias_tol, rms_tol,
                                                scratch)
        t2 = time.perf_counter()
        ok = bool(acc[0, 0])
        info = dict(predict_s=t1 - t0, check_s=t2 - t1, check_realizations=check_realizations,
                    check_bias=float(b[0, 0]), check_rms=float(rms[0, 0]), bias_tol=bias_tol,
                    rms_tol=rms_tol, accepted=ok)
        if ok:
            return q, info
        scratch.zero_()
        ctx.run_realizations(scratch, fallback_realizations, seed + 1, 0)
        full = scratch / float(fallback_realizations)
        T.cuda.synchronize()
        info["fallback_s"] = time.perf_counter() - t2
        return full, info
    finally:
        if own:
            ctx.close()
