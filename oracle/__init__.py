"""CPU oracle for the Monte Carlo drainage-basin pipeline.

TEST INFRASTRUCTURE ONLY.  A ctypes view of ``liboracle.so``, the
clean-room restatement of the reference's OpenMP CPU stages (see
``oracle.cpp`` for the reference file:line each function follows).  Only
``tests/``, the smoke check and the CPU-baseline leg of ``bench.py`` import
this module; the product (``cageo_amd``) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

EXPONENTIAL = 0
GAUSSIAN = 1
OUTSIDE = -9999
LAKE = -9998

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    c_int, c_i64, c_u64, c_u32, c_f = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                       ctypes.c_uint32, ctypes.c_float)
    lib.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    lib.or_rng_log.argtypes = [ctypes.c_double]
    lib.or_rng_log.restype = ctypes.c_double
    lib.or_rng_cossin_2pi.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
    lib.or_random_field.argtypes = [_f32p, c_i64, c_u64, c_u32, c_u32, c_f, c_f]
    lib.or_create_filter.argtypes = [c_int, c_f, c_f, c_int, _f32p]
    lib.or_create_filter.restype = c_int
    lib.or_convolve.argtypes = [_f32p, c_int, c_int, _f32p, c_int, c_int]
    lib.or_convolve.restype = c_int
    lib.or_realistic_dem.argtypes = [_f32p, _f32p, c_i64, c_f]
    lib.or_stream_burn.argtypes = [_f32p, _f32p, c_int, c_int, c_int, c_f, c_f]
    lib.or_stream_burn.restype = c_int
    lib.or_pit_fill.argtypes = [_f32p, _f32p, c_int, c_int, c_f]
    lib.or_pit_fill.restype = c_int
    lib.or_pit_fill_dijkstra.argtypes = [_f32p, _f32p, c_int, c_int, c_f]
    lib.or_pit_fill_dijkstra.restype = c_int
    lib.or_flow_routing.argtypes = [_i8p, ctypes.c_void_p, _f32p, c_int, c_int]
    lib.or_flow_tracing.argtypes = [_i32p, _i8p, _f32p, c_int, c_int]
    lib.or_flow_tracing.restype = c_i64
    lib.or_catchment_border.argtypes = [_f32p, _i32p, c_int, c_int]
    lib.or_realization_from_noise.argtypes = [
        _f32p, _f32p, _f32p, c_int, c_int, _f32p, c_int, c_int, c_f, c_f, c_int,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.or_realization_from_noise.restype = c_int
    lib.or_realization.argtypes = [_f32p, _f32p, c_int, c_int, _f32p, c_int, c_int, c_f,
                                   c_f, c_u64, c_u32, c_f, c_f, c_int, _f32p]
    lib.or_realization.restype = c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"oracle {what} failed with status {st}")


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def rng_log(u: float) -> float:
    return lib().or_rng_log(u)


def rng_cossin_2pi(t: float):
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().or_rng_cossin_2pi(t, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def random_field(shape, seed, realization=0, stream_id=0, mean=0.0, std=1.0):
    h, w = shape
    out = np.empty((h, w), np.float32)
    lib().or_random_field(out, h * w, seed, realization, stream_id, mean, std)
    return out


def create_filter(model, practical_range, cell_size, radius):
    L = 2 * radius + 1
    out = np.empty(L if model == GAUSSIAN else L * L, np.float32)
    _check(lib().or_create_filter(model, practical_range, cell_size, radius, out), "filter")
    return out if model == GAUSSIAN else out.reshape(L, L)


def filter_radius(practical_range, cell_size):
    """main.cpp:110: ceil((PR / cell) * 2.5) in float."""
    return int(np.ceil(np.float32(np.float32(practical_range) / np.float32(cell_size))
                       * np.float32(2.5)))


def convolve(field, filt, radius, model):
    f = np.ascontiguousarray(field, np.float32).copy()
    h, w = f.shape
    _check(lib().or_convolve(f, w, h, np.ascontiguousarray(filt, np.float32).ravel(),
                             radius, model), "convolve")
    return f


def realistic_dem(dem, field, nodata):
    f = np.ascontiguousarray(field, np.float32).copy()
    lib().or_realistic_dem(np.ascontiguousarray(dem, np.float32), f, f.size, nodata)
    return f


def stream_burn(z, stream, bevel_radius, nodata_dem, nodata_stream):
    z = np.ascontiguousarray(z, np.float32).copy()
    h, w = z.shape
    _check(lib().or_stream_burn(z, np.ascontiguousarray(stream, np.float32), w, h,
                                bevel_radius, nodata_dem, nodata_stream), "stream_burn")
    return z


def pit_fill(dem, nodata, method="bucket"):
    """Returns (filled, quantized_dem)."""
    d = np.ascontiguousarray(dem, np.float32).copy()
    h, w = d.shape
    spill = np.empty_like(d)
    fn = lib().or_pit_fill if method == "bucket" else lib().or_pit_fill_dijkstra
    _check(fn(spill, d, w, h, nodata), "pit_fill")
    return spill, d


def flow_routing(filled, return_flatdist=False):
    z = np.ascontiguousarray(filled, np.float32)
    h, w = z.shape
    fd = np.empty((h, w, 2), np.int8)
    flat = np.empty((h, w), np.int32) if return_flatdist else None
    lib().or_flow_routing(fd, None if flat is None else flat.ctypes.data, z, w, h)
    return (fd, flat) if return_flatdist else fd


def flow_tracing(flowdir, stream, return_guarded=False):
    fd = np.ascontiguousarray(flowdir, np.int8)
    h, w = fd.shape[:2]
    labels = np.empty((h, w), np.int32)
    g = lib().or_flow_tracing(labels, fd, np.ascontiguousarray(stream, np.float32), w, h)
    return (labels, g) if return_guarded else labels


def catchment_border(mc, labels):
    mc = np.ascontiguousarray(mc, np.float32).copy()
    h, w = mc.shape
    lib().or_catchment_border(mc, np.ascontiguousarray(labels, np.int32), w, h)
    return mc


def realization_from_noise(noise, dem, stream, filt, radius, model, nodata_dem,
                           nodata_stream, bevel_radius=2, mc=None, outputs=False):
    noise = np.ascontiguousarray(noise, np.float32).copy()
    h, w = noise.shape
    if mc is None:
        mc = np.zeros((h, w), np.float32)
    filled = np.empty((h, w), np.float32) if outputs else None
    fd = np.empty((h, w, 2), np.int8) if outputs else None
    lab = np.empty((h, w), np.int32) if outputs else None
    st = lib().or_realization_from_noise(
        noise, np.ascontiguousarray(dem, np.float32), np.ascontiguousarray(stream, np.float32),
        w, h, np.ascontiguousarray(filt, np.float32).ravel(), radius, model, nodata_dem,
        nodata_stream, bevel_radius, mc.ctypes.data,
        None if filled is None else filled.ctypes.data,
        None if fd is None else fd.ctypes.data, None if lab is None else lab.ctypes.data)
    _check(st, "realization")
    if outputs:
        return mc, filled, fd, lab
    return mc
