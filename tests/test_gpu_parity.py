"""HIP stages against the CPU oracle, through the C ABI, on the GPU.

Integer, label and byte outputs must be bit-identical; so must the float
rasters (convolution, filled DEM), because the HIP kernels round every
operation as the reference's CPU build does.  Each case prints its name,
shape and seed before touching the GPU.
"""
import numpy as np
import pytest

import oracle
from helpers import NODATA, mm_grid, plateau_grid, rng, synth_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ca(torch_cuda):
    import cageo_amd
    cageo_amd.lib()
    return cageo_amd


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def announce(*what):
    print("CASE", *what, flush=True)


# ---------------------------------------------------------------- random field
@pytest.mark.parametrize("shape", [(1000, 1024), (7, 3), (1, 1)])
def test_random_field_bit_exact(ca, torch_cuda, shape):
    announce("random_field", shape)
    torch = torch_cuda
    h, w = shape
    with ca.Context(w, h) as ctx:
        t = torch.empty((h, w), dtype=torch.float32, device="cuda")
        ctx.random_field(t, 0x1234ABCD9, 5, 3, 0.25, 1.8)
        ref = oracle.random_field(shape, 0x1234ABCD9, 5, 3, 0.25, 1.8)
        assert np.array_equal(host(t), ref)


# ----------------------------------------------------------------- convolution
# PR 260 m (r = 65): the exponential taps do not fit in LDS beside the tile
@pytest.mark.parametrize("shape,pr", [((1024, 1024), 60), ((200, 300), 60), ((31, 40), 60),
                                      ((96, 130), 200), ((256, 256), 200), ((17, 17), 30),
                                      ((150, 160), 260)])
def test_convolution_bit_exact(ca, torch_cuda, shape, pr):
    torch = torch_cuda
    h, w = shape
    for model in (oracle.GAUSSIAN, oracle.EXPONENTIAL):
        r = oracle.filter_radius(pr, 10.0)
        if r > min(h, w):
            continue
        if model == oracle.EXPONENTIAL and h * w * (2 * r + 1) ** 2 > 1e9:
            continue
        announce("convolution", shape, pr, model)
        f = ca.create_filter(model, pr, 10.0, r)
        noise = oracle.random_field(shape, 99, 1, 0, 0.0, 1.0)
        with ca.Context(w, h) as ctx:
            ctx.set_filter(model, r, f)
            t = dev(torch, noise)
            ctx.convolve(t)
            ref = oracle.convolve(noise, f, r, model)
            got = host(t)
            assert np.array_equal(got, ref), f"max diff {np.abs(got - ref).max()}"


@pytest.mark.parametrize("shape,pr", [((256, 256), 200), ((96, 130), 200), ((200, 300), 60),
                                      ((1024, 1024), 200), ((17, 17), 30), ((512, 128), 200),
                                      ((64, 2048), 30), ((2048, 1024), 260)])
def test_convolution_fast_mode(ca, torch_cuda, shape, pr):
    """Fast mode of the exponential filter: on power-of-two grids the
    periodic correlation through the FFT of the torus (conv_terms -1),
    elsewhere the low-rank separable form with fused multiply-adds (1..32
    terms).  Either is within 1e-5 of the exact convolution relative to its
    largest magnitude; the Gaussian filter, separable already, stays
    bit-exact.  The exact GPU convolution is the reference (it equals the
    oracle bit for bit, test above)."""
    torch = torch_cuda
    h, w = shape
    noise = oracle.random_field(shape, 99, 1, 0, 0.0, 1.0)
    for model in (oracle.EXPONENTIAL, oracle.GAUSSIAN):
        r = oracle.filter_radius(pr, 10.0)
        if r > min(h, w):
            continue
        announce("convolution fast mode", shape, pr, model)
        f = ca.create_filter(model, pr, 10.0, r)
        with ca.Context(w, h) as ctx:
            ctx.set_filter(model, r, f)
            t = dev(torch, noise)
            ctx.convolve(t)
            exact = host(t)
            ctx.set_conv_mode(ca.CONV_FAST)
            terms = ctx.conv_terms()
            t = dev(torch, noise)
            ctx.convolve(t)
            fast = host(t)
        if model == oracle.GAUSSIAN:
            assert terms == 0 and np.array_equal(fast, exact)
        else:
            h2, w2 = (h & (h - 1)) == 0, (w & (w - 1)) == 0
            if h2 and w2 and w >= 64 and h >= 32 and r <= min(w, h) // 2:
                assert terms == -1, terms
            else:
                assert 1 <= terms <= 32, terms
            err = np.abs(fast.astype(np.float64) - exact).max() / np.abs(exact).max()
            assert err <= 1e-5, err


def test_run_realizations_fast_mode(ca, torch_cuda):
    """A Monte Carlo run in fast mode (exponential PR 200 m, 512^2): the
    counters stay within [0, n], and their total within 1 % of the exact
    mode's (the fields differ by ~1e-6 of their range, so only cells at a
    quantisation or tie boundary can change)."""
    torch = torch_cuda
    announce("run_realizations fast mode 512^2")
    size, n = 512, 4
    dem, st = synth_case(size, seed=11)
    r = ca.filter_radius(200.0, 10.0)
    f = ca.create_filter(ca.EXPONENTIAL, 200.0, 10.0, r)
    sums = {}
    for mode in (ca.CONV_EXACT, ca.CONV_FAST):
        with ca.Context(size, size) as ctx:
            ctx.set_filter(ca.EXPONENTIAL, r, f)
            ctx.set_conv_mode(mode)
            ctx.mc_setup(dev(torch, dem), dev(torch, st), float(NODATA), float(NODATA), 2, 0.0, 2.5)
            mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
            ctx.run_realizations(mc, n, 77, 0)
            m = host(mc)
            assert m.min() >= 0 and m.max() <= n
            sums[mode] = float(m.sum())
    assert abs(sums[ca.CONV_FAST] - sums[ca.CONV_EXACT]) <= 0.01 * sums[ca.CONV_EXACT], sums


# ------------------------------------------------------------- realistic / burn
def test_realistic_dem_bit_exact(ca, torch_cuda):
    announce("realistic_dem")
    torch = torch_cuda
    r = rng(3)
    dem = r.uniform(-5, 50, (300, 257)).astype(np.float32)
    dem[r.random(dem.shape) < 0.05] = NODATA
    e = r.normal(0, 3, dem.shape).astype(np.float32)
    e[0, :5] = NODATA  # an error value equal to nodata stays
    with ca.Context(257, 300) as ctx:
        td, te = dev(torch, dem), dev(torch, e)
        ctx.realistic_dem(td, te, float(NODATA))
        assert np.array_equal(host(te), oracle.realistic_dem(dem, e, NODATA))


@pytest.mark.parametrize("bevel", [0, 1, 2, 3, 5])
def test_stream_burn_bit_exact(ca, torch_cuda, bevel):
    announce("stream_burn", bevel)
    torch = torch_cuda
    r = rng(10 + bevel)
    h, w = 97, 131
    z = r.uniform(0.5, 40, (h, w)).astype(np.float32)
    stream = np.full((h, w), NODATA, np.float32)
    stream[r.random((h, w)) < 0.08] = 1.0
    # stream cells whose DEM value is nodata exercise the raster-order rule
    z[(stream == 1.0) & (r.random((h, w)) < 0.3)] = NODATA
    z[r.random((h, w)) < 0.02] = NODATA
    with ca.Context(w, h) as ctx:
        tz = dev(torch, z)
        ctx.stream_burn(tz, dev(torch, stream), bevel, float(NODATA), float(NODATA))
        assert np.array_equal(host(tz), oracle.stream_burn(z, stream, bevel, NODATA, NODATA))


# -------------------------------------------------------------------- pit fill
def _pit_fill_case(ca, torch, z, nodata=NODATA):
    h, w = z.shape
    with ca.Context(w, h) as ctx:
        td = dev(torch, z)
        tf = torch.empty_like(td)
        ctx.pit_fill(tf, td, float(nodata))
        ref, qref = oracle.pit_fill(z, nodata)
        assert np.array_equal(host(td), qref), "quantized input differs"
        got = host(tf)
        assert np.array_equal(got, ref), f"{(got != ref).sum()} cells differ"


@pytest.mark.parametrize("seed", range(12))
def test_pit_fill_small_grids(ca, torch_cuda, seed):
    r = rng(100 + seed)
    h, w = int(r.integers(3, 150)), int(r.integers(3, 150))
    kind = seed % 4
    if kind == 0:
        z = r.uniform(0, 5, (h, w)).astype(np.float32)
    elif kind == 1:
        z = mm_grid(r, h, w, 0, 600, 0.05)
    elif kind == 2:
        z = mm_grid(r, h, w, 240, 300, 0.1)
    else:
        z = mm_grid(r, h, w, 0, 3000, 0.05, nodata_frac=0.05)
    announce("pit_fill small", seed, (h, w), kind)
    _pit_fill_case(ca, torch_cuda, z)


def test_pit_fill_synthetic_1024(ca, torch_cuda):
    announce("pit_fill synthetic 1024")
    dem, stream = synth_case(1024)
    r = oracle.filter_radius(60, 10.0)
    f = oracle.create_filter(oracle.GAUSSIAN, 60, 10.0, r)
    e = oracle.convolve(oracle.random_field(dem.shape, 5, 0, 0, 0.0, 1.8), f, r, oracle.GAUSSIAN)
    z = oracle.stream_burn(oracle.realistic_dem(dem, e, NODATA), stream, 2, NODATA, NODATA)
    _pit_fill_case(ca, torch_cuda, z)


def test_pit_fill_degenerate_shapes(ca, torch_cuda):
    r = rng(5)
    for shape in [(1, 1), (1, 9), (9, 1), (2, 2), (3, 65), (65, 3), (64, 64), (65, 129)]:
        announce("pit_fill shape", shape)
        _pit_fill_case(ca, torch_cuda, mm_grid(r, *shape, 0, 900, 0.02))
    announce("pit_fill all nodata")
    _pit_fill_case(ca, torch_cuda, np.full((20, 30), NODATA, np.float32))
    announce("pit_fill no seeds in the interior")
    z = mm_grid(r, 80, 90, 500, 900, 0.0)
    _pit_fill_case(ca, torch_cuda, z)


def test_pit_fill_rejects_negative(ca, torch_cuda):
    announce("pit_fill negative")
    torch = torch_cuda
    z = np.ones((10, 10), np.float32)
    z[4, 4] = -2.0
    with ca.Context(10, 10) as ctx:
        td = dev(torch, z)
        with pytest.raises(ca.CageoError) as ei:
            ctx.pit_fill(torch.empty_like(td), td, float(NODATA))
        assert ei.value.status == -6


# --------------------------------------------------------------- flow routing
def _routing_case(ca, torch, z):
    h, w = z.shape
    with ca.Context(w, h) as ctx:
        fd = torch.empty((h, w, 2), dtype=torch.int8, device="cuda")
        ctx.flow_routing(fd, dev(torch, z))
        ref = oracle.flow_routing(z)
        got = host(fd)
        assert np.array_equal(got, ref), f"{(got != ref).any(-1).sum()} cells differ"


@pytest.mark.parametrize("seed", range(6))
def test_flow_routing_plateaus(ca, torch_cuda, seed):
    r = rng(200 + seed)
    h, w = int(r.integers(3, 200)), int(r.integers(3, 200))
    announce("flow_routing plateaus", seed, (h, w))
    _routing_case(ca, torch_cuda, plateau_grid(r, h, w, levels=2 + seed % 3))


def test_flow_routing_filled_1024(ca, torch_cuda):
    announce("flow_routing filled 1024")
    dem, stream = synth_case(1024)
    z = oracle.stream_burn(dem, stream, 2, NODATA, NODATA)
    filled, _ = oracle.pit_fill(z, NODATA)
    _routing_case(ca, torch_cuda, filled)


# ---------------------------------------------------------------- flow tracing
def _tracing_case(ca, torch, fd, stream):
    h, w = stream.shape
    with ca.Context(w, h) as ctx:
        lab = torch.empty((h, w), dtype=torch.int32, device="cuda")
        longp = ctx.flow_tracing(lab, dev(torch, fd), dev(torch, stream))
        ref = oracle.flow_tracing(fd, stream)
        got = host(lab)
        assert longp == 0
        assert np.array_equal(got, ref), f"{(got != ref).sum()} cells differ"


@pytest.mark.parametrize("seed", range(6))
def test_flow_tracing_random_directions(ca, torch_cuda, seed):
    r = rng(300 + seed)
    h, w = int(r.integers(2, 120)), int(r.integers(2, 120))
    announce("flow_tracing random", seed, (h, w))
    fd = r.integers(-1, 2, (h, w, 2)).astype(np.int8)  # includes (0,0) and cycles
    stream = np.full((h, w), NODATA, np.float32)
    m = r.random((h, w)) < 0.05
    stream[m] = r.integers(0, 5, m.sum()).astype(np.float32)  # id 0 quirk included
    _tracing_case(ca, torch_cuda, fd, stream)


def test_flow_tracing_pipeline_1024(ca, torch_cuda):
    announce("flow_tracing pipeline 1024")
    dem, stream = synth_case(1024)
    z = oracle.stream_burn(dem, stream, 2, NODATA, NODATA)
    filled, _ = oracle.pit_fill(z, NODATA)
    _tracing_case(ca, torch_cuda, oracle.flow_routing(filled), stream)


# ------------------------------------------------------------ catchment border
def test_catchment_border_bit_exact(ca, torch_cuda):
    announce("catchment_border")
    torch = torch_cuda
    r = rng(4)
    h, w = 123, 77
    lab = r.choice(np.array([-9999, -9998, -1, -2, -3, 0], np.int32), (h, w))
    mc = r.integers(0, 5, (h, w)).astype(np.float32)
    mc[r.random((h, w)) < 0.05] = -1.0
    with ca.Context(w, h) as ctx:
        tm = dev(torch, mc)
        ctx.catchment_border(tm, dev(torch, lab))
        assert np.array_equal(host(tm), oracle.catchment_border(mc, lab))


# ------------------------------------------------------------ whole pipeline
@pytest.mark.parametrize("model,pr", [(1, 60), (0, 200)])
def test_pipeline_injected_noise_1024(ca, torch_cuda, model, pr):
    """Ten injected noise fields (seeds 1..10): filled DEM, directions,
    labels and border counts bit-identical to the oracle."""
    torch = torch_cuda
    size = 1024 if model == 1 else 256
    dem, stream = synth_case(size)
    r = oracle.filter_radius(pr, 10.0)
    f = ca.create_filter(model, pr, 10.0, r)
    with ca.Context(size, size) as ctx:
        ctx.set_filter(model, r, f)
        tdem, tst = dev(torch, dem), dev(torch, stream)
        ctx.mc_setup(tdem, tst, float(NODATA), float(NODATA), 2, 0.0, 1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        mc_ref = np.zeros((size, size), np.float32)
        seeds = range(1, 11) if model == 1 else range(1, 3)
        for s in seeds:
            announce("pipeline", model, pr, size, "seed", s)
            noise = oracle.random_field(dem.shape, s, 0, 0, 0.0, 1.8)
            filled = torch.empty((size, size), dtype=torch.float32, device="cuda")
            fd = torch.empty((size, size, 2), dtype=torch.int8, device="cuda")
            lab = torch.empty((size, size), dtype=torch.int32, device="cuda")
            ctx.realization_from_noise(dev(torch, noise), mc, filled, fd, lab)
            mc_ref, f_ref, fd_ref, lab_ref = oracle.realization_from_noise(
                noise, dem, stream, f, r, model, NODATA, NODATA, 2, mc=mc_ref, outputs=True)
            assert np.array_equal(host(filled), f_ref)
            assert np.array_equal(host(fd), fd_ref)
            assert np.array_equal(host(lab), lab_ref)
            assert np.array_equal(host(mc), mc_ref)


def test_run_realizations_matches_oracle(ca, torch_cuda):
    """Full loop including the random field: 3 realizations at 512^2."""
    torch = torch_cuda
    size = 512
    dem, stream = synth_case(size, seed=11)
    r = oracle.filter_radius(60, 10.0)
    f = ca.create_filter(1, 60, 10.0, r)
    with ca.Context(size, size) as ctx:
        ctx.set_filter(1, r, f)
        ctx.mc_setup(dev(torch, dem), dev(torch, stream), float(NODATA), float(NODATA), 2, 0.0,
                     1.8)
        mc = torch.zeros((size, size), dtype=torch.float32, device="cuda")
        announce("run_realizations 512 x3")
        ctx.run_realizations(mc, 3, 77, 10)
        ref = np.zeros((size, size), np.float32)
        for k in range(3):
            oracle.lib().or_realization(dem, stream, size, size, f, r, 1, 0.0, 1.8, 77, 10 + k,
                                        float(NODATA), float(NODATA), 2, ref)
        assert np.array_equal(host(mc), ref)
