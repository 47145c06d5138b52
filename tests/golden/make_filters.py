"""Writes tests/golden/filters.npz: process-convolution filters computed with
numpy's FFT in float64, independently of the oracle and the library.

The construction follows ConvolutionAlgorithm.cpp:18-172 of the reference:
correlogram on the (2r+1)^2 lag grid (exponential in float32 arithmetic as the
reference's float overloads, Gaussian in float64), K = shift(irfft2(|rfft2(C)|
^(1/2))) in float64, and for the Gaussian model the separable vector u0 sqrt(s0)
of numpy's SVD of the float32 kernel, sign fixed by a positive centre tap (the
reference leaves the sign to Eigen).  r = ceil(2.5 PR / cell) (main.cpp:110).

Run: python tests/golden/make_filters.py
"""
import os

import numpy as np

CASES = [(1, 60.0), (1, 200.0), (1, 30.0), (0, 60.0), (0, 200.0)]  # (model, PR); cell 10 m
CELL = 10.0


def radius(pr, cell):
    return int(np.ceil(np.float32(np.float32(pr) / np.float32(cell)) * np.float32(2.5)))


def correlogram(model, pr, cell, r):
    y, x = np.mgrid[-r:r + 1, -r:r + 1]
    d2 = (x * x + y * y).astype(np.int64)
    if model == 0:
        scale = np.float32(np.float32(pr) / np.float32(3.0)) / np.float32(cell)
        d = np.sqrt(d2.astype(np.float32))
        return np.exp((-d / scale).astype(np.float64)).astype(np.float32)
    den = np.float64(np.float32(np.float32(pr) / np.float32(1.73205080757)) / np.float32(cell)) ** 2
    return np.exp(-d2 / den).astype(np.float32)


def filter_of(model, pr, cell):
    r = radius(pr, cell)
    L = 2 * r + 1
    c = correlogram(model, pr, cell, r).astype(np.float64)
    F = np.fft.rfft2(c)
    M = np.sqrt(np.sqrt(F.real ** 2 + F.imag ** 2))
    k = np.fft.irfft2(M.astype(np.complex128), s=(L, L))
    k = np.roll(k, (L // 2, L // 2), axis=(0, 1)).astype(np.float32)
    if model == 1:
        U, S, _ = np.linalg.svd(k.astype(np.float64))
        u = U[:, 0] * np.sqrt(S[0])
        if u[r] < 0:
            u = -u
        return r, u.astype(np.float32)
    return r, k


def main():
    out = {}
    for model, pr in CASES:
        r, f = filter_of(model, pr, CELL)
        out[f"m{model}_pr{int(pr)}"] = f
        out[f"m{model}_pr{int(pr)}_radius"] = np.int32(r)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "filters.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sorted(out))


if __name__ == "__main__":
    main()
