"""Fail-safe grid barrier of the persistent kernels, forced from the host.

The persistent kernels (pit fill, flow routing, flow tracing) wait at grid
barriers.  Each wait is bounded and also ends when the host raises the
context's abort word; then every workgroup leaves normally and the call
returns CAGEO_ERR_ABORTED (-7).  These tests force both conditions from the
host (the abort word raised before the call; a wait bound of zero), so every
kernel still exits normally, and then check that the same context, cleared,
computes exactly what the oracle does.
"""
import numpy as np
import pytest

import oracle
from helpers import NODATA, mm_grid, rng, synth_case

pytestmark = pytest.mark.gpu


def announce(*what):
    print("CASE", *what, flush=True)


@pytest.fixture(scope="module")
def ca(torch_cuda):
    import cageo_amd
    cageo_amd.lib()
    return cageo_amd


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_abort_word_ends_each_persistent_stage(ca, torch_cuda):
    torch = torch_cuda
    z = mm_grid(rng(21), 300, 257, 0, 3000, 0.05)
    h, w = z.shape
    ref_fill, _ = oracle.pit_fill(z, NODATA)
    ref_fd = oracle.flow_routing(ref_fill)
    stream = np.full((h, w), NODATA, np.float32)
    stream[rng(22).random((h, w)) < 0.03] = 2.0
    ref_lab = oracle.flow_tracing(ref_fd, stream)
    with ca.Context(w, h) as ctx:
        filled = torch.empty((h, w), dtype=torch.float32, device="cuda")
        fd = torch.empty((h, w, 2), dtype=torch.int8, device="cuda")
        lab = torch.empty((h, w), dtype=torch.int32, device="cuda")
        announce("abort word raised: pit fill, routing, tracing")
        ctx.abort(True)
        for name, call in [
                ("pit_fill", lambda: ctx.pit_fill(filled, dev(torch, z), float(NODATA))),
                ("flow_routing", lambda: ctx.flow_routing(fd, dev(torch, ref_fill))),
                ("flow_tracing", lambda: ctx.flow_tracing(lab, dev(torch, ref_fd),
                                                          dev(torch, stream)))]:
            with pytest.raises(ca.CageoError) as err:
                call()
            assert err.value.status == ca.ERR_ABORTED, name
        announce("abort word cleared: the same context again")
        ctx.abort(False)
        ctx.pit_fill(filled, dev(torch, z), float(NODATA))
        assert np.array_equal(filled.cpu().numpy(), ref_fill)
        ctx.flow_routing(fd, dev(torch, ref_fill))
        assert np.array_equal(fd.cpu().numpy(), ref_fd)
        ctx.flow_tracing(lab, dev(torch, ref_fd), dev(torch, stream))
        assert np.array_equal(lab.cpu().numpy(), ref_lab)


def test_zero_wait_bound_abandons_the_realization(ca, torch_cuda):
    """With a wait bound of zero the first barrier wait gives up: the fused
    Monte Carlo call returns CAGEO_ERR_ABORTED (the launches after it end at
    once).  With the default bound back, the counters equal the oracle's."""
    torch = torch_cuda
    size = 256
    dem, stream = synth_case(size, seed=12)
    r = oracle.filter_radius(60, 10.0)
    f = ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r)
    with ca.Context(size, size) as ctx:
        ctx.set_filter(ca.GAUSSIAN, r, f)
        ctx.mc_setup(dev(torch, dem), dev(torch, stream), float(NODATA), float(NODATA), 2, 0.0,
                     1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        announce("wait bound 0: run_realizations x2")
        ctx.set_barrier_timeout(0.0)
        with pytest.raises(ca.CageoError) as err:
            ctx.run_realizations(mc, 2, 5, 0)
        assert err.value.status == ca.ERR_ABORTED
        announce("wait bound 5 s: run_realizations x2")
        ctx.set_barrier_timeout(5.0)
        mc.zero_()
        ctx.run_realizations(mc, 2, 5, 0)
        ref = np.zeros((size, size), np.float32)
        for k in range(2):
            oracle.lib().or_realization(dem, stream, size, size, f, r, 1, 0.0, 1.8, 5, k,
                                        float(NODATA), float(NODATA), 2, ref)
        assert np.array_equal(mc.cpu().numpy(), ref)


def test_abort_word_ends_all_tile_launches(ca, torch_cuda):
    """The all-tile launches of a batch read tile 0's abort word."""
    torch = torch_cuda
    shapes = [(96, 130), (70, 64)]
    tiles = []
    for k, (h, w) in enumerate(shapes):
        d, s = synth_case(max(h, w), seed=60 + k)
        tiles.append((dev(torch, d[:h, :w]), dev(torch, s[:h, :w])))
    with ca.Batch(shapes) as b:
        for t, (dd, ds) in enumerate(tiles):
            b.setup_tile(t, dd, ds, float(NODATA), float(NODATA), 2, ca.GAUSSIAN, 60.0, 10.0,
                         0.0, 1.8, 70 + t)
        mcs = [torch.zeros(s, dtype=torch.float32, device="cuda") for s in shapes]
        announce("batch: abort raised, stream mode")
        b.abort(True)
        with pytest.raises(ca.CageoError) as err:
            b.run(mcs, 1, 0, ca.BATCH_STREAM)
        assert err.value.status == ca.ERR_ABORTED
        announce("batch: abort cleared, graph vs per-tile")
        b.abort(False)
        for m in mcs:
            m.zero_()
        b.run(mcs, 2, 0, ca.BATCH_GRAPH)
        per = [torch.zeros(s, dtype=torch.float32, device="cuda") for s in shapes]
        b.run(per, 2, 0, ca.BATCH_PER_TILE)
        for a, c in zip(mcs, per):
            assert torch.equal(a, c)
