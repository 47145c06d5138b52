"""cageo_amd: MI355X (gfx950) Monte Carlo drainage-basin pipeline.

Python view of the C ABI in ``include/cageo_hip.h`` (``libcageo_hip.so``),
used by the tests and the benchmark.  Device buffers are passed as torch CUDA
tensors (only their data pointers cross the boundary) or as raw integers.
There is no CPU fallback: importing a stage without the built library raises.

Stage methods mirror the reference's stage classes (jsoksane/cageo, src/):
``random_field`` <- RandomFieldAlgorithm::execute, ``convolve`` <-
ConvolutionAlgorithm::execute, ``realistic_dem`` <- RealisticDemAlgorithm::
execute, ``stream_burn`` <- StreamBurningAlgorithm::execute, ``pit_fill`` <-
PitFillingAlgorithm::execute, ``flow_routing`` <- FlowRoutingAlgorithm::
execute, ``flow_tracing`` <- FlowTracingAlgorithm::execute,
``catchment_border`` <- CatchmentBorderAlgorithm::execute, and
``run_realizations`` <- the loop body of main.cpp:180-283.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import subprocess
import weakref

import numpy as np

__all__ = [
    "CageoError", "Context", "lib", "build", "create_filter", "filter_radius", "synth_dem",
    "EXPONENTIAL", "GAUSSIAN", "OUTSIDE", "LAKE", "LIB_PATH", "CONV_EXACT", "CONV_FAST",
    "filter_low_rank", "Surrogate",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcageo_hip.so")

EXPONENTIAL = 0
GAUSSIAN = 1
OUTSIDE = -9999
LAKE = -9998

_STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error",
           -3: "device allocation failed", -4: "iteration cap reached",
           -5: "invalid call order", -6: "input outside the supported domain",
           -7: "persistent launch abandoned (abort word or barrier time limit)"}
ERR_ABORTED = -7

# Every live handle (context, batch, surrogate), closed before the interpreter
# tears down: a handle left to the garbage collector at exit would free device
# memory after the HIP runtime or a profiler attached to it has begun its own
# teardown.
_live = weakref.WeakSet()


@atexit.register
def _close_all():
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass


class CageoError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {_STATUS.get(status, 'unknown status')} ({status})")


def build() -> None:
    """Compile libcageo_hip.so for gfx950 (hipcc, in-tree)."""
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run cageo_amd.build() (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, f, u32, u64, i64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                               ctypes.c_uint64, ctypes.c_int64)
    sigs = {
        "cageo_status_string": ([i], ctypes.c_char_p),
        "cageo_version": ([], ctypes.c_char_p),
        "cageo_context_create": ([ctypes.POINTER(vp), i, i, vp], i),
        "cageo_context_destroy": ([vp], i),
        "cageo_filter_radius": ([f, f], i),
        "cageo_create_filter": ([i, f, f, i, vp], i),
        "cageo_set_filter": ([vp, i, i, vp], i),
        "cageo_random_field": ([vp, vp, u64, u32, u32, f, f], i),
        "cageo_convolve": ([vp, vp], i),
        "cageo_realistic_dem": ([vp, vp, vp, f], i),
        "cageo_stream_burn": ([vp, vp, vp, i, f, f], i),
        "cageo_pit_fill": ([vp, vp, vp, f], i),
        "cageo_flow_routing": ([vp, vp, vp, f], i),
        "cageo_flow_tracing": ([vp, vp, vp, vp, ctypes.POINTER(i64)], i),
        "cageo_catchment_border": ([vp, vp, vp], i),
        "cageo_mc_setup": ([vp, vp, vp, f, f, i, f, f], i),
        "cageo_run_realizations": ([vp, vp, i, u64, u32], i),
        "cageo_realization_from_noise": ([vp, vp, vp, vp, vp, vp], i),
        "cageo_stage_times": ([vp, vp, i], i),
        "cageo_synth_dem": ([i, i, u64, f, i, f, vp, vp], i),
        "cageo_to_byte_array": ([vp, vp, vp], i),
        "cageo_read_header": ([ctypes.c_char_p, ctypes.POINTER(i), ctypes.POINTER(i),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(f), ctypes.POINTER(f)], i),
        "cageo_write_jpeg_gray": ([ctypes.c_char_p, vp, i, i, i], i),
        "cageo_iteration_counts": ([vp, vp], i),
        "cageo_context_configure": ([vp, i, i], i),
        "cageo_context_abort": ([vp, i], i),
        "cageo_context_set_barrier_timeout": ([vp, ctypes.c_double], i),
        "cageo_context_long_paths": ([vp, ctypes.POINTER(i64)], i),
        "cageo_batch_abort": ([vp, i], i),
        "cageo_pit_fill_noalloc": ([vp, vp, i, i, f, vp], i),
        "cageo_context_set_conv_mode": ([vp, i], i),
        "cageo_context_conv_terms": ([vp, vp], i),
        "cageo_filter_low_rank": ([vp, i, i, ctypes.c_double, vp, vp, vp, vp], i),
        "cageo_batch_create": ([ctypes.POINTER(vp), i, vp, vp, vp], i),
        "cageo_batch_destroy": ([vp], i),
        "cageo_batch_context": ([vp, i, ctypes.POINTER(vp)], i),
        "cageo_batch_setup_tile": ([vp, i, vp, vp, f, f, i, i, f, f, f, f, u64], i),
        "cageo_batch_run": ([vp, vp, i, u32, i], i),
        "cageo_batch_tune": ([vp, vp, i, vp], i),
        "cageo_batch_configure_tile": ([vp, i, i, i], i),
        "cageo_batch_info": ([vp, vp], i),
        "cageo_surrogate_create": ([ctypes.POINTER(vp), ctypes.c_char_p, vp], i),
        "cageo_surrogate_destroy": ([vp], i),
        "cageo_surrogate_forward": ([vp, vp, i, i, vp], i),
        "cageo_surrogate_predict": ([vp, vp, vp, i, i, f, f, f, f, vp], i),
        "cageo_surrogate_check": ([vp, vp, vp, i, u64, ctypes.c_double, ctypes.c_double, vp,
                                   ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)], i),
        "cageo_trainer_create": ([ctypes.POINTER(vp), i, u64, ctypes.c_char_p, vp], i),
        "cageo_trainer_destroy": ([vp], i),
        "cageo_trainer_gradients": ([vp, vp, vp, i, i, i, ctypes.POINTER(f)], i),
        "cageo_trainer_step": ([vp, vp, vp, i, i, i, f, ctypes.POINTER(f)], i),
        "cageo_surrogate_train": ([vp, vp, vp, i, i, i, i, f, f, i, i, i, u64, vp], i),
        "cageo_trainer_read": ([vp, ctypes.c_char_p, i, vp, ctypes.c_longlong], i),
        "cageo_trainer_save": ([vp, ctypes.c_char_p], i),
        "cageo_check_windows": ([vp, i, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "cageo_check_probability_map": ([vp, vp, i, u64, i, ctypes.c_double, ctypes.c_double,
                                         vp, vp, vp, vp], i),
        "cageo_refine_probability_map": ([vp, vp, i, vp, i, u64, vp], i),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(status: int, what: str) -> None:
    if status != 0:
        raise CageoError(status, what)


def _ptr(x):
    """Device pointer of a torch tensor (or an int / None)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def filter_radius(practical_range: float, cell_size: float) -> int:
    """main.cpp:110 -- ceil((PR / cell) * 2.5) in float."""
    r = lib().cageo_filter_radius(practical_range, cell_size)
    if r < 0:
        raise CageoError(r, "filter_radius")
    return r


def create_filter(model: int, practical_range: float, cell_size: float, radius: int) -> np.ndarray:
    """ConvolutionAlgorithm::createFilter (ConvolutionAlgorithm.cpp:18-45), host side."""
    L = 2 * radius + 1
    out = np.empty(L if model == GAUSSIAN else L * L, np.float32)
    _check(lib().cageo_create_filter(model, practical_range, cell_size, radius,
                                     out.ctypes.data), "create_filter")
    return out if model == GAUSSIAN else out.reshape(L, L)


def synth_dem(width: int, height: int, seed: int, cell_size: float = 10.0,
              acc_threshold: int = 2000, nodata_stream: float = -9999.0):
    """Seeded synthetic DEM and stream raster (host), as float32 numpy arrays."""
    dem = np.empty((height, width), np.float32)
    stream = np.empty((height, width), np.float32)
    _check(lib().cageo_synth_dem(width, height, seed, cell_size, acc_threshold, nodata_stream,
                                 dem.ctypes.data, stream.ctypes.data), "synth_dem")
    return dem, stream


BATCH_PER_TILE, BATCH_STREAM, BATCH_GRAPH = 0, 1, 2
CONV_EXACT, CONV_FAST = 0, 1


def filter_low_rank(taps, radius, kmax=32, tol=1e-6):
    """(values[K], vectors[K][2r+1], residual) of the symmetric filter's
    low-rank separable form (cageo_filter_low_rank, host only)."""
    n = 2 * radius + 1
    f = np.ascontiguousarray(taps, np.float32).reshape(n, n)
    vec = np.zeros((kmax, n), np.float32)
    val = np.zeros(kmax, np.float32)
    k, res = ctypes.c_int(), ctypes.c_double()
    _check(lib().cageo_filter_low_rank(f.ctypes.data, radius, kmax, tol, vec.ctypes.data,
                                       val.ctypes.data, ctypes.byref(k), ctypes.byref(res)),
           "filter_low_rank")
    return val[:k.value].copy(), vec[:k.value].copy(), res.value


class Batch:
    """Many tiles of uneven size x many realizations (cageo_batch)."""

    def __init__(self, shapes, stream=None):
        """shapes: list of (height, width)."""
        self.shapes = [tuple(map(int, s)) for s in shapes]
        n = len(self.shapes)
        ws = (ctypes.c_int * n)(*[w for _, w in self.shapes])
        hs = (ctypes.c_int * n)(*[h for h, _ in self.shapes])
        h = ctypes.c_void_p()
        sp = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream")
                                          else stream)
        _check(lib().cageo_batch_create(ctypes.byref(h), n, ws, hs, sp), "batch_create")
        self._h = h
        self._keep = []
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().cageo_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def setup_tile(self, t, dem, stream, nodata_dem, nodata_stream, bevel_radius, model,
                   practical_range, cell_size, mean, std, seed):
        self._keep.append((dem, stream))
        _check(lib().cageo_batch_setup_tile(self._h, t, _ptr(dem), _ptr(stream), nodata_dem,
                                            nodata_stream, bevel_radius, model, practical_range,
                                            cell_size, mean, std, seed), "batch_setup_tile")

    @staticmethod
    def _ptrs(mcs):
        return (ctypes.c_void_p * len(mcs))(*[_ptr(m) for m in mcs])

    def run(self, mcs, n, first=0, mode=BATCH_GRAPH):
        _check(lib().cageo_batch_run(self._h, self._ptrs(mcs), n, first, mode), "batch_run")

    def tune(self, scratch_mcs, reps=2):
        out = (ctypes.c_int * (2 * len(self.shapes)))()
        _check(lib().cageo_batch_tune(self._h, self._ptrs(scratch_mcs), reps, out), "batch_tune")
        return [(out[2 * t], out[2 * t + 1]) for t in range(len(self.shapes))]

    def info(self):
        """Blocks over all tiles and the workgroups of the all-tile launches."""
        out = (ctypes.c_int * 4)()
        _check(lib().cageo_batch_info(self._h, out), "batch_info")
        return {"blocks": out[0], "pit_fill_workgroups": out[1],
                "flow_routing_workgroups": out[2], "flow_tracing_workgroups": out[3]}

    def configure_tile(self, t, persistent_workgroups=0, conv_rows=0):
        _check(lib().cageo_batch_configure_tile(self._h, t, persistent_workgroups, conv_rows),
               "batch_configure_tile")

    def abort(self, set_=True):
        """Raise (or clear) the abort word of every tile's context."""
        _check(lib().cageo_batch_abort(self._h, int(bool(set_))), "batch_abort")

    def context(self, t):
        """The raw context handle of tile t (for stage calls on one tile)."""
        h = ctypes.c_void_p()
        _check(lib().cageo_batch_context(self._h, t, ctypes.byref(h)), "batch_context")
        return h


class Surrogate:
    """The border-probability surrogate behind the C ABI (cageo_surrogate):
    weights from a `.bin` file, inference and the Monte Carlo check on the
    device."""

    def __init__(self, weights_path, stream=None):
        h = ctypes.c_void_p()
        sp = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream")
                                          else stream)
        _check(lib().cageo_surrogate_create(ctypes.byref(h), os.fsencode(weights_path), sp),
               "surrogate_create")
        self._h = h
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().cageo_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self, x_bf16, h, w, out):
        """Prepared bf16 inputs [h][w][32] (h, w multiples of 8) -> out [h][w]."""
        _check(lib().cageo_surrogate_forward(self._h, _ptr(x_bf16), h, w, _ptr(out)),
               "surrogate_forward")

    def predict(self, dem, stream, nodata_dem, nodata_stream, rmse, practical_range, out):
        h, w = dem.shape
        _check(lib().cageo_surrogate_predict(self._h, _ptr(dem), _ptr(stream), w, h, nodata_dem,
                                             nodata_stream, rmse, practical_range, _ptr(out)),
               "surrogate_predict")

    def check(self, ctx, prob, k, seed, bias_tol, rms_tol, scratch):
        """(accepted, bias, rms) of a k-realization Monte Carlo check of the
        whole map (see Context.check_probability_map)."""
        acc, b, r = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cageo_surrogate_check(self._h, ctx._h, _ptr(prob), k, seed, bias_tol,
                                           rms_tol, _ptr(scratch), ctypes.byref(acc),
                                           ctypes.byref(b), ctypes.byref(r)), "surrogate_check")
        return bool(acc.value), b.value, r.value


class Trainer:
    """Training of the border-probability surrogate behind the C ABI
    (cageo_trainer, cageo_surrogate_train): fp32 master weights, Adam, the
    forward and backward passes on the library's BF16 MFMA kernels.
    Inputs are fp32 device tensors [tiles][h][w][4], targets [tiles][h][w]."""

    def __init__(self, base=32, seed=0, init_weights=None, stream=None):
        h = ctypes.c_void_p()
        sp = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream")
                                          else stream)
        path = None if init_weights is None else os.fsencode(init_weights)
        _check(lib().cageo_trainer_create(ctypes.byref(h), base, seed, path, sp), "trainer_create")
        self._h = h
        self.base = base
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().cageo_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def gradients(self, x4, y):
        """Loss and gradients of the batch (x4 [n][h][w][4], y [n][h][w])."""
        n, h, w = y.shape
        loss = ctypes.c_float()
        _check(lib().cageo_trainer_gradients(self._h, _ptr(x4), _ptr(y), n, h, w,
                                             ctypes.byref(loss)), "trainer_gradients")
        return loss.value

    def step(self, x4, y, lr):
        n, h, w = y.shape
        loss = ctypes.c_float()
        _check(lib().cageo_trainer_step(self._h, _ptr(x4), _ptr(y), n, h, w, lr,
                                        ctypes.byref(loss)), "trainer_step")
        return loss.value

    def train(self, x4, y, batch, lr, lr_floor, first_step, nsteps, total_steps, seed):
        """Steps first_step .. first_step + nsteps - 1 of a total_steps run
        (cageo_surrogate_train); returns their losses (numpy float32)."""
        import numpy as np
        ntiles, h, w = y.shape
        hist = np.zeros(max(nsteps, 1), np.float32)
        _check(lib().cageo_surrogate_train(self._h, _ptr(x4), _ptr(y), ntiles, h, w, batch, lr,
                                           lr_floor, first_step, nsteps, total_steps, seed,
                                           hist.ctypes.data), "surrogate_train")
        return hist[:nsteps]

    def read(self, name, shape, grad=False):
        import numpy as np
        out = np.zeros(shape, np.float32)
        _check(lib().cageo_trainer_read(self._h, name.encode(), 1 if grad else 0,
                                        out.ctypes.data, out.size), "trainer_read")
        return out

    def save(self, path):
        _check(lib().cageo_trainer_save(self._h, os.fsencode(path)), "trainer_save")


def pit_fill_noalloc(filled, dem, nodata_dem, stream=None) -> None:
    """cageo_pit_fill_noalloc: the pit fill with no context and no device
    allocation (workspace in the library's own device array)."""
    h, w = dem.shape
    sp = None if stream is None else (stream.cuda_stream if hasattr(stream, "cuda_stream")
                                      else stream)
    _check(lib().cageo_pit_fill_noalloc(_ptr(filled), _ptr(dem), w, h, nodata_dem, sp),
           "pit_fill_noalloc")


def read_header(path: str) -> dict:
    """The .hdr raster header (utils.cpp:58-112 loadHeaderData)."""
    c, r = ctypes.c_int(), ctypes.c_int()
    x, y = ctypes.c_double(), ctypes.c_double()
    cs, nd = ctypes.c_float(), ctypes.c_float()
    _check(lib().cageo_read_header(os.fsencode(path), ctypes.byref(c), ctypes.byref(r),
                                   ctypes.byref(x), ctypes.byref(y), ctypes.byref(cs),
                                   ctypes.byref(nd)), f"read_header({path})")
    return {"ncols": c.value, "nrows": r.value, "xllcorner": x.value, "yllcorner": y.value,
            "cellsize": cs.value, "NODATA_value": nd.value}


def write_jpeg_gray(path: str, pixels: np.ndarray, quality: int = 90) -> None:
    """Baseline grayscale JPEG of a uint8 H x W image (the CLI's .jpg output)."""
    px = np.ascontiguousarray(pixels, dtype=np.uint8)
    h, w = px.shape
    _check(lib().cageo_write_jpeg_gray(os.fsencode(path), px.ctypes.data, w, h, quality),
           "write_jpeg_gray")


class Context:
    """Device workspace for one width x height grid (cageo_context)."""

    def __init__(self, width: int, height: int, stream=None):
        self.width, self.height = width, height
        h = ctypes.c_void_p()
        s = stream if isinstance(stream, (int, type(None))) else stream.cuda_stream
        _check(lib().cageo_context_create(ctypes.byref(h), width, height, s), "context_create")
        self._h = h
        _live.add(self)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().cageo_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stages -------------------------------------------------------------
    def set_filter(self, model: int, radius: int, taps) -> None:
        t = np.ascontiguousarray(taps, np.float32).ravel()
        _check(lib().cageo_set_filter(self._h, model, radius, t.ctypes.data), "set_filter")

    def random_field(self, field, seed, realization=0, stream_id=0, mean=0.0, std=1.0):
        _check(lib().cageo_random_field(self._h, _ptr(field), seed, realization, stream_id,
                                        mean, std), "random_field")

    def convolve(self, field):
        _check(lib().cageo_convolve(self._h, _ptr(field)), "convolve")

    def realistic_dem(self, dem, field, nodata_dem):
        _check(lib().cageo_realistic_dem(self._h, _ptr(dem), _ptr(field), nodata_dem),
               "realistic_dem")

    def stream_burn(self, z, stream, bevel_radius, nodata_dem, nodata_stream):
        _check(lib().cageo_stream_burn(self._h, _ptr(z), _ptr(stream), bevel_radius,
                                       nodata_dem, nodata_stream), "stream_burn")

    def pit_fill(self, filled, dem, nodata_dem):
        _check(lib().cageo_pit_fill(self._h, _ptr(filled), _ptr(dem), nodata_dem), "pit_fill")

    def flow_routing(self, flowdir, filled, nodata_dem=-9999.0):
        _check(lib().cageo_flow_routing(self._h, _ptr(flowdir), _ptr(filled), nodata_dem),
               "flow_routing")

    def flow_tracing(self, labels, flowdir, stream) -> int:
        lp = ctypes.c_int64(0)
        _check(lib().cageo_flow_tracing(self._h, _ptr(labels), _ptr(flowdir), _ptr(stream),
                                        ctypes.byref(lp)), "flow_tracing")
        return lp.value

    def catchment_border(self, mc, labels):
        _check(lib().cageo_catchment_border(self._h, _ptr(mc), _ptr(labels)), "catchment_border")

    # -- Monte Carlo driver -------------------------------------------------
    def mc_setup(self, dem, stream, nodata_dem, nodata_stream, bevel_radius=2, mean=0.0,
                 std=1.0):
        self._inputs = (dem, stream)  # keep the device buffers alive
        _check(lib().cageo_mc_setup(self._h, _ptr(dem), _ptr(stream), nodata_dem, nodata_stream,
                                    bevel_radius, mean, std), "mc_setup")

    def run_realizations(self, mc, n, seed, first=0):
        _check(lib().cageo_run_realizations(self._h, _ptr(mc), n, seed, first),
               "run_realizations")

    def realization_from_noise(self, noise, mc=None, filled=None, flowdir=None, labels=None):
        _check(lib().cageo_realization_from_noise(self._h, _ptr(noise), _ptr(mc), _ptr(filled),
                                                  _ptr(flowdir), _ptr(labels)),
               "realization_from_noise")

    def to_byte_array(self, out_bytes, values):
        _check(lib().cageo_to_byte_array(self._h, _ptr(out_bytes), _ptr(values)), "to_byte_array")

    def iteration_counts(self):
        """(pit-fill rounds, pit-fill passes, flat passes, tracing rounds) of the last run."""
        out = np.zeros(4, np.int32)
        _check(lib().cageo_iteration_counts(self._h, out.ctypes.data), "iteration_counts")
        return tuple(int(v) for v in out)

    def set_conv_mode(self, mode):
        """CONV_EXACT (bit-exact reference order) or CONV_FAST (low-rank form of
        the exponential filter, within 1e-5)."""
        _check(lib().cageo_context_set_conv_mode(self._h, mode), "set_conv_mode")

    def conv_terms(self):
        out = ctypes.c_int()
        _check(lib().cageo_context_conv_terms(self._h, ctypes.byref(out)), "conv_terms")
        return out.value

    def configure(self, persistent_workgroups=0, conv_rows=0):
        _check(lib().cageo_context_configure(self._h, persistent_workgroups, conv_rows),
               "context_configure")

    def check_windows(self, window=0):
        """(windows along x, windows along y) of a window size."""
        nx, ny = ctypes.c_int(), ctypes.c_int()
        _check(lib().cageo_check_windows(self._h, window, ctypes.byref(nx), ctypes.byref(ny)),
               "check_windows")
        return nx.value, ny.value

    def check_probability_map(self, prob, k, seed, bias_tol, rms_tol, scratch, window=0):
        """Monte Carlo check of a probability map with k realizations of this
        context: per window (row-major; window 0 = the whole raster) the
        unbiased mean signed error and RMS error estimates and the verdict
        |bias| <= bias_tol and rms <= rms_tol.  Returns numpy arrays
        (accepted, bias, rms) of shape (windows along y, windows along x)."""
        nx, ny = self.check_windows(window)
        acc = np.zeros(nx * ny, np.int32)
        b = np.zeros(nx * ny, np.float64)
        r = np.zeros(nx * ny, np.float64)
        _check(lib().cageo_check_probability_map(self._h, _ptr(prob), k, seed, window, bias_tol,
                                                 rms_tol, _ptr(scratch), acc.ctypes.data,
                                                 b.ctypes.data, r.ctypes.data),
               "check_probability_map")
        return acc.reshape(ny, nx).astype(bool), b.reshape(ny, nx), r.reshape(ny, nx)

    def refine_probability_map(self, prob, accepted, n, seed, scratch, window=0):
        """Replace the windows of `prob` whose flag in `accepted` is False by
        the frequency of an n-realization Monte Carlo of the whole DEM."""
        acc = np.ascontiguousarray(np.asarray(accepted).ravel(), np.int32)
        _check(lib().cageo_refine_probability_map(self._h, _ptr(prob), window, acc.ctypes.data,
                                                  n, seed, _ptr(scratch)),
               "refine_probability_map")

    def abort(self, set_=True):
        """Raise (or clear) the abort word: while it is raised, persistent
        launches end at their start, or in a barrier wait that stalls, with
        CAGEO_ERR_ABORTED."""
        _check(lib().cageo_context_abort(self._h, int(bool(set_))), "context_abort")

    def set_barrier_timeout(self, seconds):
        _check(lib().cageo_context_set_barrier_timeout(self._h, float(seconds)),
               "set_barrier_timeout")

    def long_paths(self):
        """Cells past the reference's walk guard, summed over the last call."""
        out = ctypes.c_int64()
        _check(lib().cageo_context_long_paths(self._h, ctypes.byref(out)), "long_paths")
        return out.value

    def stage_times(self, enable=True):
        out = np.zeros(7, np.float32)
        _check(lib().cageo_stage_times(self._h, out.ctypes.data, int(enable)), "stage_times")
        return out
