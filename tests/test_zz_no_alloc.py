"""The allocation-free pit fill (cageo_pit_fill_noalloc) against the
context entry point (cageo_pit_fill), through the C ABI, on the GPU.

Two calls in a row on one device share the library's device array: the
second, smaller one must not see the first one's contents.  Each result must
be bit-identical to cageo_pit_fill on the same input (and to the oracle).
(A grid beyond the array's bound is refused on the host before any device
work; tests/test_library_cpu.py covers that without a GPU.)
"""
import numpy as np
import pytest

import oracle
from helpers import NODATA, mm_grid, rng, synth_case

pytestmark = pytest.mark.gpu


def announce(*what):
    print("CASE", *what, flush=True)


def _both(ca, torch, z):
    h, w = z.shape
    with ca.Context(w, h) as ctx:
        a = torch.from_numpy(z.copy()).cuda()
        fa = torch.empty_like(a)
        ctx.pit_fill(fa, a, float(NODATA))
        ref = fa.cpu().numpy()
    b = torch.from_numpy(z.copy()).cuda()
    fb = torch.empty_like(b)
    ca.pit_fill_noalloc(fb, b, float(NODATA))
    torch.cuda.synchronize()
    return ref, fb.cpu().numpy(), b.cpu().numpy()


def test_noalloc_matches_context_entry_point(torch_cuda):
    import cageo_amd as ca
    torch = torch_cuda
    dem, stream = synth_case(1024, seed=9)
    big = oracle.stream_burn(dem, stream, 2, NODATA, NODATA)
    small = mm_grid(rng(31), 300, 257, 0, 3000, 0.05, nodata_frac=0.02)
    for name, z in (("synthetic 1024^2 burned", big), ("mm grid 300 x 257", small)):
        announce("pit_fill_noalloc", name)
        ref, got, quant = _both(ca, torch, z)
        assert np.array_equal(got, ref), f"{name}: {int((got != ref).sum())} cells differ"
        oref, oq = oracle.pit_fill(z, NODATA)
        assert np.array_equal(quant, oq) and np.array_equal(got, oref)
