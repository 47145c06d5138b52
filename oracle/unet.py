"""Float64 restatement of the border-probability surrogate's forward and
backward pass.  Test infrastructure only: it checks the gradients of the
library's BF16 MFMA kernels (the network has no counterpart in the reference,
so the mathematics is the specification: 3x3 zero-padded convolutions over
channel concatenations with nearest 2x upsampling, ReLU, 2x2 average
pooling, a 1x1 head and the mean binary cross-entropy of sigmoid outputs).

Layout and parameter names follow cageo_amd.surrogate: activations NHWC,
convolution weights [9][out][in] with tap = 3*(dy+1) + (dx+1).
"""
from __future__ import annotations

import numpy as np


def _shift(x, dy, dx):
    """x[n, y+dy, x+dx, c] with zeros outside (NHWC)."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, h), slice(0, h - dy)) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
    xs, xd = (slice(dx, w), slice(0, w - dx)) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up2_back(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool2_back(g):
    return up2(g) * 0.25


def conv3x3(x, w, b):
    """y[p][o] = b[o] + sum_tap sum_i x[p + off(tap)][i] w[tap][o][i]."""
    y = np.zeros(x.shape[:3] + (w.shape[1],))
    for tap in range(9):
        y += _shift(x, tap // 3 - 1, tap % 3 - 1) @ w[tap].T
    return y + b


def conv3x3_back(x, w, g):
    """Gradients (dx, dw, db) of conv3x3 for output gradient g."""
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    gf = g.reshape(-1, g.shape[-1])
    for tap in range(9):
        dy, ddx = tap // 3 - 1, tap % 3 - 1
        xs = _shift(x, dy, ddx)
        dw[tap] = gf.T @ xs.reshape(-1, x.shape[-1])
        dx += _shift(g @ w[tap], -dy, -ddx)
    return dx, dw, g.sum(axis=(0, 1, 2))


SPEC = ["enc1a", "enc1b", "enc2a", "enc2b", "enc3a", "enc3b", "mida", "midb",
        "dec3a", "dec3b", "dec2a", "dec2b", "dec1a", "dec1b"]


def forward_backward(params, x, target, rnd=None):
    """Loss and gradients (dict like params) for inputs x [N][H][W][C0] and
    target frequencies [N][H][W]; float64 arithmetic.  `rnd`, when given, is
    applied where a reduced-precision implementation stores values: every
    layer output (after ReLU), every pooled tensor and every gradient that
    enters a convolution's backward pass."""
    P = {k: np.asarray(v, np.float64) for k, v in params.items()}
    x = np.asarray(x, np.float64)
    t = np.asarray(target, np.float64)
    R = rnd if rnd is not None else (lambda a: a)
    cache = {}

    def layer(name, inp):
        pre = conv3x3(inp, P[name + ".w"], P[name + ".b"])
        out = R(np.maximum(pre, 0.0))
        cache[name] = (inp, pre)
        return out

    def pool(a):
        return R(pool2(a))

    e1 = layer("enc1b", layer("enc1a", x))
    e2 = layer("enc2b", layer("enc2a", pool(e1)))
    e3 = layer("enc3b", layer("enc3a", pool(e2)))
    mid = layer("midb", layer("mida", pool(e3)))
    d3 = layer("dec3b", layer("dec3a", np.concatenate([up2(mid), e3], axis=-1)))
    d2 = layer("dec2b", layer("dec2a", np.concatenate([up2(d3), e2], axis=-1)))
    d1 = layer("dec1b", layer("dec1a", np.concatenate([up2(d2), e1], axis=-1)))
    z = d1 @ P["head.w"] + P["head.b"][0]
    q = 1.0 / (1.0 + np.exp(-z))
    npx = z.size
    loss = float(np.mean(np.maximum(z, 0) - z * t + np.log1p(np.exp(-np.abs(z)))))
    grads = {}
    gz = (q - t) / npx
    grads["head.w"] = np.einsum("nhwc,nhw->c", d1, gz)
    grads["head.b"] = np.array([gz.sum()])
    g = gz[..., None] * P["head.w"]

    def back(name, g_out):
        inp, pre = cache[name]
        g_pre = R(g_out * (pre > 0))
        dx, dw, db = conv3x3_back(inp, P[name + ".w"], g_pre)
        grads[name + ".w"], grads[name + ".b"] = dw, db
        return dx

    def split(dcat, c_up):
        return up2_back(dcat[..., :c_up]), dcat[..., c_up:]

    d = back("dec1a", back("dec1b", g))
    g_d2, g_e1 = split(d, d2.shape[-1])
    d = back("dec2a", back("dec2b", g_d2))
    g_d3, g_e2 = split(d, d3.shape[-1])
    d = back("dec3a", back("dec3b", g_d3))
    g_mid, g_e3 = split(d, mid.shape[-1])
    d = back("mida", back("midb", g_mid))
    g_e3 = g_e3 + pool2_back(d)
    d = back("enc3a", back("enc3b", g_e3))
    g_e2 = g_e2 + pool2_back(d)
    d = back("enc2a", back("enc2b", g_e2))
    g_e1 = g_e1 + pool2_back(d)
    back("enc1a", back("enc1b", g_e1))
    return loss, grads, q


def loss_only(params, x, target):
    return forward_backward(params, x, target)[0]


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), kept in float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
