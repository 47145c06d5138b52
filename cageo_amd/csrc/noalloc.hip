// Allocation-free pit fill: cageo_pit_fill with its workspace in a file-scope
// device array instead of a context's allocations, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).  It runs the same kernels as
// cageo_pit_fill (k_reset_minmax, k_quantize_minmax, k_pf_persistent,
// k_pf_resident) on the same workspace layout, carved from the array.
//
// The array is a __device__ symbol: every device that loads the library holds
// its own copy, at its own address, until the process ends.  Its host-side
// name is only a stand-in, so the device address is taken once per device
// with hipGetSymbolAddress (that device current) and used for every kernel
// argument.  All calls on one device share that device's copy and leave the
// last call's contents; each call clears what must start at zero (the tile
// epoch stamps, the control block and the barrier words) with a fill kernel.
// Calls that overlap on one device race on the array: callers serialise them.
#include <mutex>

#include "common.hpp"

namespace cageo {

namespace {

constexpr int64_t NA_MAX_CELLS = CAGEO_NOALLOC_MAX_CELLS;
constexpr int64_t NA_MAX_TILES = CAGEO_NOALLOC_MAX_TILES;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
// Workspace layout (bytes): the context buffers cageo_pit_fill uses.
constexpr size_t OFF_ST = 0;                                             // cell states
constexpr size_t OFF_WORK = align256(OFF_ST + NA_MAX_CELLS);             // 7 tile lists
constexpr size_t OFF_MARK = align256(OFF_WORK + 7 * NA_MAX_TILES * 4);   // epoch stamps
constexpr size_t OFF_TOPEN = align256(OFF_MARK + NA_MAX_TILES * 4);      // open flags
constexpr size_t OFF_TCOST = align256(OFF_TOPEN + NA_MAX_TILES * 4);     // solve costs
constexpr size_t OFF_ROWS = align256(OFF_TCOST + NA_MAX_TILES * 4);      // dirty-row masks
constexpr size_t OFF_CTRL = align256(OFF_ROWS + 4 * NA_MAX_TILES * 4);   // control block
constexpr size_t OFF_BAR = align256(OFF_CTRL + CTL_WORDS * 4);           // barrier words
constexpr size_t OFF_SCAL = align256(OFF_BAR + BAR_WORDS * 4);           // min / max / bad
constexpr size_t SCRATCH_BYTES = align256(OFF_SCAL + SC_NSLOTS * 4);
static_assert(SCRATCH_BYTES <= CAGEO_NOALLOC_SCRATCH_BYTES, "the declared bound holds the layout");

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// Per device: the array's device address and the launch settings.
struct DeviceScratch {
    uint8_t* ptr = nullptr;
    int pf_grid = 0, pf_res_grid = 0;
    uint64_t bar_limit = 0;
    bool ready = false;
};
constexpr int NA_MAX_DEVICES = 64;
DeviceScratch g_dev[NA_MAX_DEVICES];
std::mutex g_dev_mu;

int device_scratch(int device, DeviceScratch** out) {
    if (device < 0 || device >= NA_MAX_DEVICES) return CAGEO_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_dev_mu);
    DeviceScratch& d = g_dev[device];
    if (!d.ready) {
        void* p = nullptr;
        CAGEO_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return CAGEO_ERR_HIP;
        Context probe;  // host-side settings only: no device memory
        probe.device = device;
        CAGEO_TRY(pit_fill_grid_cap(&probe, &d.pf_grid));
        CAGEO_TRY(pit_fill_resident_cap(&probe, &d.pf_res_grid));
        CAGEO_TRY(barrier_ticks(device, BAR_DEFAULT_TIMEOUT_S, &d.bar_limit));
        d.ptr = static_cast<uint8_t*>(p);
        d.ready = true;
    }
    *out = &d;
    return CAGEO_OK;
}

// The words that must start at zero: epoch stamps of `ntiles` tiles, the
// control block, the barrier block.
__global__ void k_noalloc_clear(uint32_t* __restrict__ mark, int ntiles,
                                uint32_t* __restrict__ ctrl, uint32_t* __restrict__ bar) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ntiles; i += gridDim.x * blockDim.x)
        mark[i] = 0u;
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < CTL_WORDS; i += blockDim.x) ctrl[i] = 0u;
        for (int i = threadIdx.x; i < BAR_WORDS; i += blockDim.x) bar[i] = 0u;
    }
}

}  // namespace

}  // namespace cageo

extern "C" int cageo_pit_fill_noalloc(float* d_filled, float* d_dem, int width, int height,
                                      float nodata_dem, void* hip_stream) {
    using namespace cageo;
    if (!d_filled || !d_dem || d_filled == d_dem || width <= 0 || height <= 0)
        return CAGEO_ERR_ARG;
    const int64_t n = (int64_t)width * height;
    const int ntx = (width + TILE - 1) / TILE, nty = (height + TILE - 1) / TILE;
    if (n > NA_MAX_CELLS || (int64_t)ntx * nty > NA_MAX_TILES) {
        std::fprintf(stderr,
                     "cageo_pit_fill_noalloc: %d x %d exceeds the workspace (%lld cells, %lld "
                     "64x64 tiles at most)\n",
                     width, height, (long long)NA_MAX_CELLS, (long long)NA_MAX_TILES);
        return CAGEO_ERR_ARG;
    }
    int device = 0;
    CAGEO_HIP_TRY(hipGetDevice(&device));
    DeviceScratch* d = nullptr;
    CAGEO_TRY(device_scratch(device, &d));

    Context c;  // a view of the array laid out as a context's pit-fill buffers
    c.W = width;
    c.H = height;
    c.n = (size_t)n;
    c.device = device;
    c.stream = reinterpret_cast<hipStream_t>(hip_stream);
    c.ntx = ntx;
    c.nty = nty;
    c.ntiles = ntx * nty;
    c.d_st = d->ptr + OFF_ST;
    c.d_work3 = reinterpret_cast<int*>(d->ptr + OFF_WORK);
    c.d_mark_pf = reinterpret_cast<uint32_t*>(d->ptr + OFF_MARK);
    c.d_topen = reinterpret_cast<uint32_t*>(d->ptr + OFF_TOPEN);
    c.d_tcost = reinterpret_cast<uint32_t*>(d->ptr + OFF_TCOST);
    c.d_rows = reinterpret_cast<uint32_t*>(d->ptr + OFF_ROWS);
    c.d_ctrl = reinterpret_cast<uint32_t*>(d->ptr + OFF_CTRL);
    c.d_bar = reinterpret_cast<uint32_t*>(d->ptr + OFF_BAR);
    c.d_scal = reinterpret_cast<uint32_t*>(d->ptr + OFF_SCAL);
    c.d_habort = nullptr;  // no abort word; the barrier waits stay bounded
    c.bar_limit = d->bar_limit;
    c.pf_grid = d->pf_grid;
    c.pf_res_grid = d->pf_res_grid;

    const int cgrid = std::max(1, std::min(64, (c.ntiles + 255) / 256));
    hipLaunchKernelGGL(k_noalloc_clear, dim3(cgrid), dim3(256), 0, c.stream, c.d_mark_pf,
                       c.ntiles, c.d_ctrl, c.d_bar);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(reset_minmax(&c));
    CAGEO_TRY(launch_quantize(&c, d_dem, nodata_dem));
    CAGEO_TRY(pit_fill_quantized(&c, d_dem, d_filled, nodata_dem, true));
    uint32_t st = 0;
    CAGEO_HIP_TRY(hipMemcpyAsync(&st, c.d_ctrl + CTL_STATUS, sizeof(st), hipMemcpyDeviceToHost,
                                 c.stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c.stream));
    return st ? -(int)st : CAGEO_OK;
}
