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
