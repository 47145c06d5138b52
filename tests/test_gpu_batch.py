"""Batched form on the GPU: tiles of uneven size x realizations.

Every batch mode (per-tile blocking calls, one stream, a replayed hipGraph)
and any autotuned setting must give the counter rasters that each tile gets
from cageo_run_realizations alone, which in turn equal the oracle's Monte
Carlo loop.  Each case prints its name before it touches the GPU.
"""
import numpy as np
import pytest

import oracle
from helpers import NODATA, synth_case

pytestmark = pytest.mark.gpu

SHAPES = [(96, 200), (257, 130), (64, 64), (190, 333), (128, 80)]
PARAMS = [  # (model, practical range, error std)
    (1, 60.0, 1.8), (0, 200.0, 2.5), (1, 30.0, 0.5), (1, 200.0, 3.0), (0, 60.0, 1.0)]


@pytest.fixture(scope="module")
def tiles(torch_cuda):
    torch = torch_cuda
    out = []
    for k, (h, w) in enumerate(SHAPES):
        dem, st = synth_case(max(h, w), seed=40 + k)
        dem, st = np.ascontiguousarray(dem[:h, :w]), np.ascontiguousarray(st[:h, :w])
        out.append((dem, st, torch.from_numpy(dem).cuda(), torch.from_numpy(st).cuda()))
    return out


def announce(*what):
    print("CASE", *what, flush=True)


def make_batch(ca, tiles):
    b = ca.Batch(SHAPES)
    for t, ((model, pr, std), (_, _, ddem, dst)) in enumerate(zip(PARAMS, tiles)):
        # radius must fit the tile: PR 200 m -> r = 50 <= 64
        b.setup_tile(t, ddem, dst, float(NODATA), float(NODATA), 2, model, pr, 10.0, 0.0, std,
                     1000 + t)
    return b


def zeros(torch):
    return [torch.zeros(s, dtype=torch.float32, device="cuda") for s in SHAPES]


def test_batch_modes_agree_with_single_tiles_and_oracle(torch_cuda, tiles):
    import cageo_amd as ca
    torch = torch_cuda
    n = 3
    announce("batch per-tile/stream/graph", SHAPES)
    with make_batch(ca, tiles) as b:
        results = {}
        for mode in (ca.BATCH_PER_TILE, ca.BATCH_STREAM, ca.BATCH_GRAPH):
            mcs = zeros(torch)
            b.run(mcs, n, first=5, mode=mode)
            results[mode] = [m.cpu().numpy() for m in mcs]
        # a second graph replay series continues the realization index
        mcs = zeros(torch)
        b.run(mcs, 1, first=5, mode=ca.BATCH_GRAPH)
        b.run(mcs, 2, first=6, mode=ca.BATCH_GRAPH)
        results["graph_split"] = [m.cpu().numpy() for m in mcs]
    for t, ((model, pr, std), (dem, st, ddem, dst)) in enumerate(zip(PARAMS, tiles)):
        h, w = SHAPES[t]
        r = ca.filter_radius(pr, 10.0)
        taps = ca.create_filter(model, pr, 10.0, r)
        with ca.Context(w, h) as ctx:
            ctx.set_filter(model, r, taps)
            ctx.mc_setup(ddem, dst, float(NODATA), float(NODATA), 2, 0.0, std)
            mc = torch.zeros((h, w), dtype=torch.float32, device="cuda")
            ctx.run_realizations(mc, n, 1000 + t, 5)
            single = mc.cpu().numpy()
        ref = np.zeros((h, w), np.float32)
        for k in range(n):
            oracle.lib().or_realization(dem, st, w, h, taps, r, model, 0.0, std, 1000 + t, 5 + k,
                                        float(NODATA), float(NODATA), 2, ref)
        assert np.array_equal(single, ref), f"tile {t}"
        for mode, res in results.items():
            assert np.array_equal(res[t], single), f"tile {t} mode {mode}"


def test_batch_tuned_settings_do_not_change_results(torch_cuda, tiles):
    import cageo_amd as ca
    torch = torch_cuda
    announce("batch tune")
    with make_batch(ca, tiles) as b:
        base = zeros(torch)
        b.run(base, 2, first=0, mode=ca.BATCH_STREAM)
        cfg = b.tune(zeros(torch), reps=1)
        assert len(cfg) == len(SHAPES) and all(g >= 1 for g, _ in cfg)
        tuned = zeros(torch)
        b.run(tuned, 2, first=0, mode=ca.BATCH_GRAPH)
        for a, c in zip(base, tuned):
            assert torch.equal(a, c)
        # hand-set extremes: a single workgroup and the full chip
        for t in range(len(SHAPES)):
            b.configure_tile(t, persistent_workgroups=1 if t % 2 else 0, conv_rows=4 if t == 1 else 0)
        other = zeros(torch)
        b.run(other, 2, first=0, mode=ca.BATCH_STREAM)
        for a, c in zip(base, other):
            assert torch.equal(a, c)


def test_batch_tile_outside_domain(torch_cuda, tiles):
    """A tile whose elevation span exceeds the reference's bucket range
    (2^24 mm, PitFillingAlgorithmCPUOpenMP.h:111-118) fails with
    CAGEO_ERR_RANGE, alone: in the all-tile launches the other tile's counter
    raster still equals its single-tile result."""
    import cageo_amd as ca
    torch = torch_cuda
    announce("batch tile outside the pit-fill domain")
    dem0, st0, _, dst0 = tiles[0]
    bad = dem0.copy()
    bad[10, 10] = 17000.0  # span > 16777.216 m
    dbad = torch.from_numpy(bad).cuda()
    _, _, ddem2, dst2 = tiles[2]
    shapes = [SHAPES[0], SHAPES[2]]
    (m0, pr0, s0), (m2, pr2, s2) = PARAMS[0], PARAMS[2]
    with ca.Batch(shapes) as b:
        b.setup_tile(0, dbad, dst0, float(NODATA), float(NODATA), 2, m0, pr0, 10.0, 0.0, s0, 7)
        b.setup_tile(1, ddem2, dst2, float(NODATA), float(NODATA), 2, m2, pr2, 10.0, 0.0, s2, 8)
        for mode in (ca.BATCH_STREAM, ca.BATCH_GRAPH):
            mcs = [torch.zeros(s, dtype=torch.float32, device="cuda") for s in shapes]
            with pytest.raises(ca.CageoError) as err:
                b.run(mcs, 2, first=0, mode=mode)
            assert err.value.status == -6
            h, w = SHAPES[2]
            r = ca.filter_radius(pr2, 10.0)
            with ca.Context(w, h) as ctx:
                ctx.set_filter(m2, r, ca.create_filter(m2, pr2, 10.0, r))
                ctx.mc_setup(ddem2, dst2, float(NODATA), float(NODATA), 2, 0.0, s2)
                single = torch.zeros((h, w), dtype=torch.float32, device="cuda")
                ctx.run_realizations(single, 2, 8, 0)
            assert torch.equal(mcs[1], single), mode


def test_batch_64_uneven_tiles_graph_vs_oracle(torch_cuda):
    """The batched workload: 64 tiles with sides drawn U(256, 2048), error
    RMSE U(0.5, 3) m and practical range U(30, 200) m (Gaussian filter), as
    the benchmark draws them; two realizations replayed as one hipGraph of
    all-tile launches.  Every tile's counter raster must equal the oracle's
    two realizations of that tile."""
    import cageo_amd as ca
    from concurrent.futures import ThreadPoolExecutor
    from helpers import oracle_counts_parallel
    torch = torch_cuda
    seed = 20240601
    r = np.random.default_rng(seed + 3)
    n = 64
    shapes = [tuple(int(v) for v in r.integers(256, 2049, size=2)) for _ in range(n)]
    rmse = r.uniform(0.5, 3.0, n)
    pr = r.uniform(30.0, 200.0, n)
    announce("batch 64 tiles U(256,2048): generating DEMs")
    with ThreadPoolExecutor(16) as ex:
        dems = list(ex.map(lambda k: ca.synth_dem(shapes[k][1], shapes[k][0], seed + 100 + k,
                                                  10.0, 2000, float(NODATA)), range(n)))
    reals = 2
    announce("batch 64 tiles U(256,2048): graph mode x2, cells", sum(h * w for h, w in shapes))
    with ca.Batch(shapes) as b:
        dev = [(torch.from_numpy(d).cuda(), torch.from_numpy(s).cuda()) for d, s in dems]
        for t, (dd, ds) in enumerate(dev):
            b.setup_tile(t, dd, ds, float(NODATA), float(NODATA), 2, ca.GAUSSIAN, float(pr[t]),
                         10.0, 0.0, float(rmse[t]), seed + 1000 + t)
        mcs = [torch.zeros(s, dtype=torch.float32, device="cuda") for s in shapes]
        b.run(mcs, reals, 0, ca.BATCH_GRAPH)
        got = [m.cpu().numpy() for m in mcs]
        info = b.info()
        print("all-tile launch", info, flush=True)
        assert info["blocks"] == sum(((h + 63) // 64) * ((w + 63) // 64) for h, w in shapes)
    announce("batch 64 tiles: oracle in worker processes")
    jobs = []
    for t in range(n):
        rad = ca.filter_radius(float(pr[t]), 10.0)
        taps = ca.create_filter(ca.GAUSSIAN, float(pr[t]), 10.0, rad)
        jobs.append((dems[t][0], dems[t][1], taps, rad, ca.GAUSSIAN, float(rmse[t]),
                     seed + 1000 + t, 0, reals))
    refs = oracle_counts_parallel(jobs)
    for t in range(n):
        assert np.array_equal(got[t], refs[t]), \
            f"tile {t} {shapes[t]}: {int((got[t] != refs[t]).sum())} counters differ"
