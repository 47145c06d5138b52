"""The 4096 x 4096 headline workload against the oracle, on the GPU.

The benchmark's DEM (synthetic 4096^2, 10 m cells, seed 20240601) with
injected noise fields: ten with the Gaussian filter of practical range 60 m
(seeds 1..10) and three with the exponential filter of practical range
200 m (seeds 1..3), error std 1.8 m.  For every field the filled DEM, the D8
directions, the basin labels and the border counter must be bit-identical to
the oracle's (the reference's CPU stages, restated).  At this size the pit
fill runs its full multi-round schedule over 4096 tiles (two-ended pass
lists, claims, phase lists and the compacted fill list), which smaller cases
do not reach.  A last case runs two full realizations (random field
included) through the fused loop against the oracle's loop.

Each case prints its name and seed before it touches the GPU.
"""
import numpy as np
import pytest

import oracle
from helpers import NODATA

pytestmark = pytest.mark.gpu

SIZE = 4096
DEM_SEED = 20240601


def announce(*what):
    print("CASE", *what, flush=True)


@pytest.fixture(scope="module")
def ca(torch_cuda):
    import cageo_amd
    cageo_amd.lib()
    return cageo_amd


@pytest.fixture(scope="module")
def dem4096(ca, torch_cuda):
    dem, stream = ca.synth_dem(SIZE, SIZE, DEM_SEED, 10.0, 2000, float(NODATA))
    torch = torch_cuda
    return dem, stream, torch.from_numpy(dem).cuda(), torch.from_numpy(stream).cuda()


def _field_case(ca, torch, dem4096, model, pr, seed):
    dem, stream, tdem, tst = dem4096
    r = oracle.filter_radius(pr, 10.0)
    f = ca.create_filter(model, pr, 10.0, r)
    noise = oracle.random_field(dem.shape, seed, 0, 0, 0.0, 1.8)
    with ca.Context(SIZE, SIZE) as ctx:
        ctx.set_filter(model, r, f)
        ctx.mc_setup(tdem, tst, float(NODATA), float(NODATA), 2, 0.0, 1.8)
        mc = torch.zeros((SIZE, SIZE), dtype=torch.float32, device="cuda")
        filled = torch.empty((SIZE, SIZE), dtype=torch.float32, device="cuda")
        fd = torch.empty((SIZE, SIZE, 2), dtype=torch.int8, device="cuda")
        lab = torch.empty((SIZE, SIZE), dtype=torch.int32, device="cuda")
        ctx.realization_from_noise(torch.from_numpy(noise).cuda(), mc, filled, fd, lab)
        rounds, passes, flat_passes, tr_rounds = ctx.iteration_counts()
        print(f"pit fill {rounds} rounds / {passes} passes, {flat_passes} flat passes, "
              f"{tr_rounds} tracing rounds", flush=True)
        assert ctx.long_paths() == 0
        got = (filled.cpu().numpy(), fd.cpu().numpy(), lab.cpu().numpy(), mc.cpu().numpy())
    mc_ref, f_ref, fd_ref, lab_ref = oracle.realization_from_noise(
        noise, dem, stream, f, r, model, NODATA, NODATA, 2, outputs=True)
    names = ("filled DEM", "D8 directions", "labels", "border counter")
    for name, g, e in zip(names, got, (f_ref, fd_ref, lab_ref, mc_ref)):
        bad = (g != e).reshape(SIZE, SIZE, -1).any(-1)
        assert not bad.any(), f"{name}: {int(bad.sum())} cells differ, first at " \
                              f"{tuple(int(v) for v in np.argwhere(bad)[0])}"
    assert mc_ref.sum() > 0


@pytest.mark.parametrize("seed", range(1, 11))
def test_gaussian_pr60_injected_field(ca, torch_cuda, dem4096, seed):
    announce("4096^2 Gaussian PR 60 m, injected field seed", seed)
    _field_case(ca, torch_cuda, dem4096, oracle.GAUSSIAN, 60.0, seed)


@pytest.mark.parametrize("seed", range(1, 4))
def test_exponential_pr200_injected_field(ca, torch_cuda, dem4096, seed):
    announce("4096^2 exponential PR 200 m, injected field seed", seed)
    _field_case(ca, torch_cuda, dem4096, oracle.EXPONENTIAL, 200.0, seed)


def test_run_realizations_4096(ca, torch_cuda, dem4096):
    """Two realizations of the fused loop (random field included) against
    the oracle's loop, counters accumulated."""
    torch = torch_cuda
    dem, stream, tdem, tst = dem4096
    r = oracle.filter_radius(60, 10.0)
    f = ca.create_filter(ca.GAUSSIAN, 60.0, 10.0, r)
    announce("4096^2 run_realizations x2, seed", DEM_SEED)
    with ca.Context(SIZE, SIZE) as ctx:
        ctx.set_filter(ca.GAUSSIAN, r, f)
        ctx.mc_setup(tdem, tst, float(NODATA), float(NODATA), 2, 0.0, 1.8)
        mc = torch.zeros((SIZE, SIZE), dtype=torch.float32, device="cuda")
        ctx.run_realizations(mc, 2, DEM_SEED, 0)
        assert ctx.long_paths() == 0
        got = mc.cpu().numpy()
    ref = np.zeros((SIZE, SIZE), np.float32)
    for k in range(2):
        st = oracle.lib().or_realization(dem, stream, SIZE, SIZE, f, r, 1, 0.0, 1.8, DEM_SEED, k,
                                         float(NODATA), float(NODATA), 2, ref)
        assert st == 0
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} counters differ"
