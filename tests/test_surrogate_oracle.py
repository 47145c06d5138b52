"""Pins the float64 surrogate oracle (oracle/unet.py) on the CPU.

- conv3x3 against a direct 7-loop evaluation of the definition;
- adjointness: <A x, y> == <x, A^T y> for the convolution, the 2x2 average
  pool and the nearest 2x upsampling (the backward pass uses A^T);
- central finite differences of the full loss against the analytic gradient
  for weights and biases spread over every layer.
These fix the gradient the GPU kernels are checked against.
"""
import numpy as np
import pytest

from cageo_amd.surrogate import init_params
from oracle import unet


def test_conv3x3_definition():
    r = np.random.default_rng(0)
    x = r.normal(size=(2, 5, 4, 3))
    w = r.normal(size=(9, 2, 3))
    b = r.normal(size=2)
    y = unet.conv3x3(x, w, b)
    ref = np.zeros((2, 5, 4, 2))
    for n in range(2):
        for yy in range(5):
            for xx in range(4):
                for o in range(2):
                    s = b[o]
                    for tap in range(9):
                        sy, sx = yy + tap // 3 - 1, xx + tap % 3 - 1
                        if 0 <= sy < 5 and 0 <= sx < 4:
                            s += (x[n, sy, sx, :] * w[tap, o, :]).sum()
                    ref[n, yy, xx, o] = s
    assert np.abs(y - ref).max() < 1e-12


def test_adjoints():
    r = np.random.default_rng(1)
    x = r.normal(size=(2, 6, 8, 3))
    w = r.normal(size=(9, 4, 3))
    g = r.normal(size=(2, 6, 8, 4))
    dx, dw, _ = unet.conv3x3_back(x, w, g)
    assert abs((unet.conv3x3(x, w, 0.0) * g).sum() - (x * dx).sum()) < 1e-9
    assert abs((unet.conv3x3(x, w, 0.0) * g).sum() - (w * dw).sum()) < 1e-9
    gp = r.normal(size=(2, 3, 4, 3))
    assert abs((unet.pool2(x) * gp).sum() - (x * unet.pool2_back(gp)).sum()) < 1e-12
    xs = r.normal(size=(2, 3, 4, 3))
    assert abs((unet.up2(xs) * x).sum() - (xs * unet.up2_back(x)).sum()) < 1e-12


def test_finite_differences():
    params = {k: v.astype(np.float64) for k, v in init_params(base=2, seed=3).items()}
    r = np.random.default_rng(4)
    x = np.zeros((1, 8, 8, 32))
    x[..., :4] = r.normal(size=(1, 8, 8, 4))
    t = r.uniform(0, 1, (1, 8, 8))
    for k in params:
        if k.endswith(".b"):
            params[k] = params[k] + r.normal(0, 0.1, params[k].shape)  # avoid ReLU kinks at 0
    _, grads, _ = unet.forward_backward(params, x, t)
    h = 1e-6
    checked = 0
    for k in sorted(params):
        flat = params[k].reshape(-1)
        idx = [i for i in r.choice(flat.size, size=min(3, flat.size), replace=False)
               if not (k == "enc1a.w" and i % 32 >= 4)]  # padded input channels see zeros
        for i in idx:
            old = flat[i]
            flat[i] = old + h
            lp = unet.loss_only(params, x, t)
            flat[i] = old - h
            lm = unet.loss_only(params, x, t)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            an = grads[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1e-3, abs(an)) + 1e-9, (k, i, fd, an)
            checked += 1
    assert checked > 40


def test_expected_l1_of_sample():
    """The check's noise level E|B/k - q|, B ~ Binomial(k, q): closed form
    2q(1-q) at k = 1, zero at q in {0, 1}, and brute-force sums at k = 3."""
    from cageo_amd.surrogate import expected_l1_of_sample
    q = np.array([0.0, 0.1, 0.5, 0.73, 1.0])
    for qi in q:
        assert abs(expected_l1_of_sample(np.array([qi]), 1) - 2 * qi * (1 - qi)) < 1e-12
    assert expected_l1_of_sample(np.array([0.0, 1.0]), 8) == 0.0
    import itertools
    for qi in q:
        e = 0.0
        for draws in itertools.product([0, 1], repeat=3):
            pr = np.prod([qi if d else 1 - qi for d in draws])
            e += pr * abs(sum(draws) / 3 - qi)
        assert abs(expected_l1_of_sample(np.array([qi]), 3) - e) < 1e-12
    assert abs(expected_l1_of_sample(q, 3) - np.mean(
        [expected_l1_of_sample(np.array([qi]), 3) for qi in q])) < 1e-12


def test_make_inputs_planes():
    """Input planes: z-scored DEM with nodata cells 0, the 0/1 stream mask,
    RMSE / 3 and practical range / 200 as constant planes, zero padding."""
    from cageo_amd.surrogate import make_inputs
    dem = np.array([[1.0, 2.0], [3.0, -9999.0]], np.float32)
    x = make_inputs(dem, dem > 2.5, 1.5, 100.0, -9999.0, channels=6)
    mu, sd = 2.0, np.std([1.0, 2.0, 3.0])
    assert np.allclose(x[..., 0], [[(1 - mu) / sd, 0.0], [(3 - mu) / sd, 0.0]])
    assert np.array_equal(x[..., 1], [[0, 0], [1, 0]])
    assert np.all(x[..., 2] == 0.5) and np.all(x[..., 3] == 0.5) and np.all(x[..., 4:] == 0)
