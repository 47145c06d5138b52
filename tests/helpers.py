"""Shared input builders for the parity tests (seeded, deterministic)."""
import numpy as np

NODATA = np.float32(-9999.0)


def rng(seed):
    return np.random.default_rng(seed)


def synth_case(size, seed=7):
    """Synthetic DEM + stream raster from the library's host generator."""
    import cageo_amd as ca
    return ca.synth_dem(size, size, seed, 10.0, 2000, float(NODATA))


def mm_grid(r, h, w, lo=0, hi=600, zero_frac=0.1, nodata_frac=0.0):
    """Elevations on whole millimetres (exercises dead levels such as 0.251)."""
    z = (r.integers(lo, hi, (h, w)) / 1000.0).astype(np.float32)
    z[r.random((h, w)) < zero_frac] = 0.0
    if nodata_frac:
        z[r.random((h, w)) < nodata_frac] = NODATA
    return z


def plateau_grid(r, h, w, levels=4):
    """Integer terraces: many exact ties and flats."""
    z = r.integers(1, levels + 1, (h, w)).astype(np.float32)
    z[r.random((h, w)) < 0.02] = 0.0
    return z


# ---- the oracle's Monte Carlo counters of many tiles, in worker processes ----
def _omp_single_thread():
    import os
    os.environ["OMP_NUM_THREADS"] = "1"  # read when the oracle library loads


def oracle_counts(job):
    """Counter raster of `n` oracle realizations of one tile (a worker)."""
    import oracle
    dem, stream, taps, r, model, std, seed, first, n = job
    h, w = dem.shape
    mc = np.zeros((h, w), np.float32)
    for k in range(n):
        st = oracle.lib().or_realization(dem, stream, w, h, taps, r, model, 0.0, std, seed,
                                         first + k, float(NODATA), float(NODATA), 2, mc)
        if st != 0:
            raise RuntimeError(f"oracle realization failed with status {st}")
    return mc


def oracle_counts_parallel(jobs, workers=16):
    """oracle_counts of every job, in spawned single-threaded worker
    processes (at most 16: the host share of one GPU), largest tiles first."""
    import multiprocessing
    import os
    from concurrent.futures import ProcessPoolExecutor
    try:
        workers = max(1, min(workers, len(os.sched_getaffinity(0))))
    except AttributeError:
        pass
    order = sorted(range(len(jobs)), key=lambda k: -jobs[k][0].size)
    out = [None] * len(jobs)
    with ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("spawn"),
                             initializer=_omp_single_thread) as ex:
        for k, mc in zip(order, ex.map(oracle_counts, [jobs[k] for k in order])):
            out[k] = mc
    return out
