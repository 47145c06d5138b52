"""Pins the oracle's counter-based normal generator (CPU only).

Philox4x32-10 against the published known-answer vectors of Random123
(Salmon et al., SC'11); the double-precision log / sincos against the C math
library; the normal draws by their distribution.  The reference's own stream
(thrust minstd / cuRAND XORWOW seeded with time()) is not reproducible and is
not a parity target.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert list(oracle.philox4x32_10(ctr, key)) == expect


def test_log_against_libm():
    us = [2.0 ** -24, 1e-6, 0.1, 0.25, 0.3, 0.5, 0.7071067811865476, 0.75, 0.999999, 1.0]
    us += list(np.random.default_rng(0).uniform(1e-7, 1.0, 2000))
    for u in us:
        assert abs(oracle.rng_log(u) - math.log(u)) <= 4e-16 * max(1.0, abs(math.log(u)))


def test_cossin_against_libm():
    for t in list(np.arange(0, 1, 1 / 64)) + list(np.random.default_rng(1).uniform(0, 1, 2000)):
        c, s = oracle.rng_cossin_2pi(t)
        assert abs(c - math.cos(2 * math.pi * t)) < 1e-15
        assert abs(s - math.sin(2 * math.pi * t)) < 1e-15


def test_normal_distribution():
    x = oracle.random_field((512, 512), 20240601, 3, 0, 0.0, 1.0).ravel().astype(np.float64)
    assert abs(x.mean()) < 0.01
    assert abs(x.std() - 1.0) < 0.01
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # lag-1 autocorrelation along the raster and between the two draws of a pair
    assert abs(np.corrcoef(x[:-1], x[1:])[0, 1]) < 0.01
    assert abs(np.corrcoef(x[0::2], x[1::2])[0, 1]) < 0.01


def test_mean_std_scaling_and_counters():
    a = oracle.random_field((64, 64), 7, 0, 0, 0.0, 1.0)
    b = oracle.random_field((64, 64), 7, 0, 0, 2.5, 1.8)
    assert np.array_equal(b, np.float32(2.5) + a * np.float32(1.8))
    # realization, stream id and seed each select an independent stream
    c = oracle.random_field((64, 64), 7, 1, 0, 0.0, 1.0)
    d = oracle.random_field((64, 64), 7, 0, 1, 0.0, 1.0)
    e = oracle.random_field((64, 64), 8, 0, 0, 0.0, 1.0)
    for other in (c, d, e):
        assert abs(np.corrcoef(a.ravel(), other.ravel())[0, 1]) < 0.1
    # deterministic and independent of the raster shape (cell i <- block i/4)
    f = oracle.random_field((32, 128), 7, 0, 0, 0.0, 1.0)
    assert np.array_equal(a.ravel(), f.ravel())
