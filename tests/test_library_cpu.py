"""The HIP library on the CPU side: it loads, exports every function the C
header declares, and its host-only entry points (filter construction, status
strings, synthetic inputs) behave.  No device calls are made here."""
import os
import re

import numpy as np
import pytest

import cageo_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cageo_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cageo_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_entry_points():
    names = declared_functions()
    for must in ["cageo_context_create", "cageo_pit_fill", "cageo_flow_routing",
                 "cageo_flow_tracing", "cageo_catchment_border", "cageo_run_realizations"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ca.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_status_strings_and_version():
    lib = ca.lib()
    assert lib.cageo_status_string(0) == b"ok"
    assert lib.cageo_status_string(-6) == b"input outside the supported domain"
    assert b"gfx950" in lib.cageo_version()


def test_create_filter_argument_checks():
    lib = ca.lib()
    out = np.zeros(31, np.float32)
    assert lib.cageo_create_filter(7, 60.0, 10.0, 15, out.ctypes.data) == -1
    assert lib.cageo_create_filter(1, 60.0, 0.0, 15, out.ctypes.data) == -1
    assert lib.cageo_filter_radius(60.0, 0.0) == -1


def test_synthetic_inputs():
    dem, st = ca.synth_dem(256, 192, 5, 10.0, 400, -9999.0)
    dem2, st2 = ca.synth_dem(256, 192, 5, 10.0, 400, -9999.0)
    assert np.array_equal(dem, dem2) and np.array_equal(st, st2)
    assert dem.shape == (192, 256) and dem.dtype == np.float32
    assert 90.0 < dem.min() and dem.max() < 100.0 + 0.05 * 256 + 40 * 60 + 5
    ids = st[st != -9999.0]
    assert ids.size > 0 and ids.min() >= 1.0 and np.all(ids == np.round(ids))
    # tilt of 0.05 m per cell along +x (dominates the hills at 1024 cells)
    big, _ = ca.synth_dem(1024, 1024, 5, 10.0, 2000, -9999.0)
    assert big[:, 512:].mean() - big[:, :512].mean() > 10.0
    dem3, _ = ca.synth_dem(256, 192, 6, 10.0, 400, -9999.0)
    assert not np.array_equal(dem, dem3)


def test_context_requires_device_or_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present; covered by the GPU tests")
    with pytest.raises(ca.CageoError):
        ca.Context(16, 16)


@pytest.mark.parametrize("pr", [30.0, 60.0, 200.0])
def test_filter_low_rank_against_eigh(pr):
    """The fast convolution's low-rank form (host Jacobi eigen-decomposition)
    against numpy's symmetric eigen-solver: the same leading eigenvalues,
    orthonormal vectors, and the reconstruction within the tolerance."""
    r = ca.filter_radius(pr, 10.0)
    f = ca.create_filter(ca.EXPONENTIAL, pr, 10.0, r).astype(np.float64)
    lam, vec, res = ca.filter_low_rank(f, r, kmax=32, tol=1e-6)
    k = len(lam)
    assert 1 <= k <= 32 and res <= 1e-6
    w = np.linalg.eigvalsh(0.5 * (f + f.T))
    w = w[np.argsort(-np.abs(w))][:k]
    assert np.allclose(lam, w, rtol=1e-5, atol=1e-7 * np.abs(w[0]))
    v = vec.astype(np.float64)
    assert np.allclose(v @ v.T, np.eye(k), atol=1e-5)
    rec = (v.T * lam) @ v
    assert np.linalg.norm(rec - f) <= 2e-6 * np.linalg.norm(f)


def test_filter_low_rank_refuses_asymmetric_and_short_budgets():
    r = 8
    f = ca.create_filter(ca.EXPONENTIAL, 30.0, 10.0, r).astype(np.float32)
    g = f.copy()
    g[0, 1] += 1e-3 * f.max()
    with pytest.raises(ca.CageoError):
        ca.filter_low_rank(g, r)
    with pytest.raises(ca.CageoError):  # two terms cannot reach 1e-6
        ca.filter_low_rank(f, r, kmax=2)


def test_pit_fill_noalloc_refuses_grids_beyond_its_workspace():
    """cageo_pit_fill_noalloc checks the grid against its fixed workspace
    before any device call: too many cells, or too many 64 x 64 tiles for a
    skinny grid, give CAGEO_ERR_ARG (the pointers are never touched)."""
    lib = ca.lib()
    for w, h in ((4097, 4096), (1 << 21, 8), (0, 10), (10, -1)):
        assert lib.cageo_pit_fill_noalloc(16, 32, w, h, -9999.0, None) == -1, (w, h)
    assert lib.cageo_pit_fill_noalloc(16, 16, 8, 8, -9999.0, None) == -1  # filled aliases dem
