"""Pins the oracle's convolution, realistic-DEM and stream-burning stages (CPU).

Convolution (ConvolutionAlgorithmCPUOpenMP.h:70-248): a periodic correlation,
checked by a delta input (returns the point-reflected filter around the
delta), a constant input (c times the tap sum), and a float64 brute force.
Realistic DEM (RealisticDemAlgorithm.h:49-70) and burning
(StreamBurningAlgorithmCPUOpenMP.h:32-82) by closed forms on hand grids.
"""
import numpy as np
import pytest

import oracle

NODATA = np.float32(-9999.0)


def brute_correlate(x, f, r):
    """float64 periodic correlation out[y,x] = sum f[a,b] in[y-r+a, x-r+b]."""
    h, w = x.shape
    out = np.zeros((h, w))
    L = 2 * r + 1
    for a in range(L):
        for b in range(L):
            out += f[a, b] * np.roll(np.roll(x, r - a, axis=0), r - b, axis=1)
    return out


@pytest.mark.parametrize("model,pr,shape", [(1, 60, (40, 37)), (0, 60, (33, 41)), (1, 30, (17, 17))])
def test_delta_and_constant(model, pr, shape):
    r = oracle.filter_radius(pr, 10.0)
    f = oracle.create_filter(model, pr, 10.0, r)
    f2 = np.outer(f, f) if model == oracle.GAUSSIAN else f
    h, w = shape
    d = np.zeros(shape, np.float32)
    y0, x0 = 3, w - 2  # near the wrap in both directions
    d[y0, x0] = 1.0
    out = oracle.convolve(d, f, r, model)
    # out[y, x] = f2[y0 - y + r, x0 - x + r] (indices mod the grid)
    exp = np.zeros(shape)
    for a in range(2 * r + 1):
        for b in range(2 * r + 1):
            exp[(y0 - a + r) % h, (x0 - b + r) % w] += f2[a, b]
    assert np.abs(out - exp).max() < 1e-6
    c = np.full(shape, 2.5, np.float32)
    outc = oracle.convolve(c, f, r, model)
    assert np.abs(outc - 2.5 * f2.sum()).max() < 1e-4 * 2.5 * abs(f2.sum())


@pytest.mark.parametrize("model", [oracle.GAUSSIAN, oracle.EXPONENTIAL])
def test_against_float64_brute_force(model):
    r = oracle.filter_radius(60, 10.0)
    f = oracle.create_filter(model, 60, 10.0, r)
    f2 = np.outer(f, f).astype(np.float64) if model == oracle.GAUSSIAN else f.astype(np.float64)
    x = oracle.random_field((48, 52), 3, 0, 0, 0.0, 1.0)
    got = oracle.convolve(x, f, r, model).astype(np.float64)
    ref = brute_correlate(x.astype(np.float64), f2, r)
    # norm-wise bound: every cell sums 2r+1 (or (2r+1)^2) float products
    scale = brute_correlate(np.abs(x.astype(np.float64)), np.abs(f2), r)
    assert (np.abs(got - ref) <= 1e-5 * scale).all()


def test_radius_larger_than_grid_rejected():
    f = oracle.create_filter(oracle.GAUSSIAN, 60, 10.0, 15)
    with pytest.raises(RuntimeError):
        oracle.convolve(np.zeros((10, 40), np.float32), f, 15, oracle.GAUSSIAN)


def test_realistic_dem_rules():
    dem = np.array([[10.0, NODATA, -5.0, 0.5]], np.float32)
    e = np.array([[1.0, 3.25, 2.0, -0.75]], np.float32)
    out = oracle.realistic_dem(dem, e, NODATA)
    # dem+e; nodata DEM keeps the raw error (output aliases the error field);
    # -5+2 = -3 <= 0 -> 1.0; 0.5-0.75 <= 0 -> 1.0
    assert out.tolist() == [[11.0, 3.25, 1.0, 1.0]]
    e2 = np.array([[1.0, -1.5, 2.0, 1.0]], np.float32)
    assert oracle.realistic_dem(dem, e2, NODATA).tolist() == [[11.0, 1.0, 1.0, 1.5]]


def test_burn_factors_and_order():
    # one stream cell in the middle of a 7x7 plateau at 8 m
    z = np.full((7, 7), 8.0, np.float32)
    s = np.full((7, 7), NODATA, np.float32)
    s[3, 3] = 1.0
    out = oracle.stream_burn(z, s, 2, NODATA, NODATA)
    assert out[3, 3] == 0.0
    f = np.float32(np.sqrt(np.float32(2.0)) / np.float32(2.0))
    for (dy, dx), fac in {(0, 1): 0.5, (1, 0): 0.5, (1, 1): f, (-1, -1): f, (2, 0): 1.0,
                          (0, -2): 1.0}.items():
        assert out[3 + dy, 3 + dx] == np.float32(8.0) * np.float32(fac)
    # beyond the radius (d^2 = 5, 8) untouched
    assert out[3 + 1, 3 + 2] == 8.0 and out[3 + 2, 3 + 2] == 8.0
    # nearest stream wins; equal distance -> order (d^2, y, x)
    z2 = np.full((5, 5), 4.0, np.float32)
    s2 = np.full((5, 5), NODATA, np.float32)
    s2[0, 2] = 1.0
    s2[2, 4] = 1.0
    out2 = oracle.stream_burn(z2, s2, 2, NODATA, NODATA)
    assert out2[1, 3] == np.float32(4.0) * f  # both diagonal at d^2 = 2
    assert out2[2, 2] == 4.0  # both at d^2 = 4 -> factor 1.0
    assert out2[1, 2] == 2.0  # (0,-1) at d^2 = 1


def test_burn_raster_order_rule():
    # A stream cell whose DEM value is nodata counts as a bevel source only for
    # cells after it in raster order: the reference zeroes it in place when
    # its turn comes, and the cells before it still read nodata.
    z = np.full((3, 5), 6.0, np.float32)
    z[1, 2] = NODATA
    s = np.full((3, 5), NODATA, np.float32)
    s[1, 2] = 1.0
    out = oracle.stream_burn(z, s, 2, NODATA, NODATA)
    f = np.float32(np.sqrt(np.float32(2.0)) / np.float32(2.0))
    assert out[1, 2] == 0.0
    for y, x in [(0, 1), (0, 2), (0, 3), (1, 0), (1, 1)]:  # before: no source
        assert out[y, x] == 6.0, (y, x)
    assert out[1, 3] == 3.0 and out[2, 2] == 3.0  # after, d^2 = 1
    assert out[2, 1] == np.float32(6.0) * f and out[2, 3] == np.float32(6.0) * f
    assert out[1, 4] == 6.0  # after, d^2 = 4: factor 1


def test_burn_radius_limits():
    z = np.ones((4, 4), np.float32)
    s = np.full((4, 4), NODATA, np.float32)
    with pytest.raises(RuntimeError):
        oracle.stream_burn(z, s, 6, NODATA, NODATA)
    assert np.array_equal(oracle.stream_burn(z, s, 0, NODATA, NODATA), z)
