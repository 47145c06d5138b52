// Shared definitions of the gfx950 library: the context (device workspace),
// launch helpers and small device utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "cageo_hip.h"

namespace cageo {

// Tiles of the iterative stencil kernels: 64 x 64 cells, 256 threads, each
// thread owning a vertical strip of 16 consecutive cells of one column so a
// wavefront reads 64 consecutive LDS words per step.
constexpr int TILE = 64;
constexpr int TILE_THREADS = 256;
constexpr int STRIP = TILE * TILE / TILE_THREADS;  // 16
constexpr int HALO_W = TILE + 2;                   // tile + 1-cell halo

// Pit-fill cell states.
enum : uint8_t { PF_NODATA = 0, PF_FIXED_LIVE = 1, PF_FIXED_DEAD = 2, PF_OPEN = 3 };

// Stream classification bits (precomputed once per run from the stream raster).
enum : uint8_t { SF_BURN = 1, SF_TRACE = 2 };

// Scalars kept in device memory (one 64-word block, read back in one copy).
enum ScalarSlot : int {
    SC_MIN = 0,        // ordered-uint min of the quantized DEM (incl. nodata)
    SC_MAX = 1,        // ordered-uint max
    SC_BAD = 2,        // negative / NaN elevation seen
    SC_COUNT0 = 3,     // worklist count, buffer 0
    SC_COUNT1 = 4,     // worklist count, buffer 1
    SC_CERTIFIED = 5,  // cells certified in this round
    SC_OPEN = 6,       // OPEN cells left
    SC_CHANGED = 7,    // generic "something changed" flag
    SC_LONGPATH = 8,   // cells whose trace is longer than the walk guard
    SC_EPOCH = 9,      // unused on device; reserved
    SC_LONGRUN = 10,   // SC_LONGPATH summed over the launches since the last clear_status
    SC_NSLOTS = 64
};

struct Context {
    int W = 0, H = 0;
    size_t n = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    // tiles
    int ntx = 0, nty = 0, ntiles = 0;
    uint32_t* d_mark = nullptr;  // per-tile epoch stamps of the flat-resolution worklists
    uint32_t* d_mark_pf = nullptr;  // epoch stamps of the pit-fill kernel
    uint32_t* d_topen = nullptr;    // per tile: holds OPEN pit-fill cells
    uint32_t* d_tcost = nullptr;    // per tile: row steps of its last pit-fill solve
    uint32_t* d_rows = nullptr;     // per tile: dirty-row masks, 2 pass parities x 64 bits
    int* d_work3 = nullptr;         // 7 worklists of the persistent kernels
    uint32_t* d_ctrl = nullptr;     // control block of the persistent kernels
    uint32_t* d_bar = nullptr;      // grid barrier words (BAR_WORDS), see grid_sync
    uint32_t* h_abort = nullptr;    // abort word in host-mapped memory (cageo_context_abort)
    uint32_t* d_habort = nullptr;   // its device address
    uint64_t bar_limit = 0;         // bound of one barrier wait, wall_clock64() ticks
    double bar_timeout_s = 0.0;
    int pf_grid = 0, fr_grid = 0, tr_grid = 0;
    int pf_res_grid = 0;  // co-resident workgroups of the pit fill's resident rounds
    int conv_ty = 0;  // tuned row-block height of the 2-D convolution (0 = largest that fits)
    // Fast mode of the 2-D convolution: the filter's low-rank separable form.
    int conv_fast = 0;
    int lr_k = 0;                     // terms
    float* d_lr_col = nullptr;        // K x LC column taps
    float* d_lr_row = nullptr;        // K x LP row taps (lambda_k v_k)
    float* d_lr_planes = nullptr;     // K column-pass planes of W x H
    std::vector<float> h_taps;        // the 2-D filter (host copy)
    // ... or its FFT form on power-of-two grids (convfft.hip)
    bool fft_ready = false;
    float2* d_fft_spec = nullptr;     // H x (W/2 + 1) half spectra
    float* d_fft_q = nullptr;         // (W/2 + 1) x (H/2 + 1) filter spectrum (scaled)
    float2* d_fft_tw = nullptr;       // exp(-2 pi i k / max(W, H))

    // scalars
    uint32_t* d_scal = nullptr;
    uint32_t* d_bmm = nullptr;   // per-block (min, max, bad) of k_realistic_burn_quant
    uint32_t* h_scal = nullptr;  // pinned mirror

    // filter
    int model = -1, radius = 0, filter_len = 0;
    float* d_filter = nullptr;

    // bound Monte Carlo inputs
    const float* dem = nullptr;
    const float* streamr = nullptr;
    float nodata_dem = 0.0f, nodata_stream = 0.0f, mean = 0.0f, std_dev = 1.0f;
    int bevel = 2;
    bool mc_ready = false;
    uint8_t* d_sflags = nullptr;
    int32_t* d_slabel = nullptr;

    // burn offsets (x, y, factor), ordered as the reference's search array
    int n_offsets = 0;
    int* d_off = nullptr;      // 2 ints per offset
    float* d_factor = nullptr;  // one factor per offset
    int off_radius = -1;

    // work rasters
    float* d_a = nullptr;  // noise / error field
    float* d_b = nullptr;  // convolution buffer
    float* d_z = nullptr;  // quantized burned DEM
    float* d_w = nullptr;  // filled DEM
    uint8_t* d_st = nullptr;
    int8_t* d_fd = nullptr;
    int32_t* d_dist = nullptr;
    int32_t* d_lab = nullptr;
    uint64_t* d_ptr = nullptr;
    uint64_t* d_bent = nullptr;  // flow tracing: 256 border-table slots per tile

    // timing
    bool timing = false;
    hipEvent_t ev[8] = {};
    std::vector<hipEvent_t> evpool;  // 8 per timed realization, recorded without host syncs
    float stage_ms[7] = {};
};

#define CAGEO_HIP_TRY(expr)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            std::fprintf(stderr, "cageo: %s failed: %s (%s:%d)\n", #expr,          \
                         hipGetErrorString(_e), __FILE__, __LINE__);               \
            return CAGEO_ERR_HIP;                                                  \
        }                                                                          \
    } while (0)

#define CAGEO_TRY(expr)             \
    do {                            \
        int _s = (expr);            \
        if (_s != CAGEO_OK) return _s; \
    } while (0)

// Float <-> order-preserving uint (for atomic min/max of floats).
__host__ __device__ inline uint32_t float_to_ordered(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ordered_to_float(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __builtin_bit_cast(float, b);
}

// toIndex / toElev of PitFillingAlgorithmCPUOpenMP.h:55-65 (PREC = 1000).
__host__ __device__ inline uint64_t pf_to_index(float elev, float min_elev) {
    return (uint64_t)(1000.0f * (elev - min_elev));
}
__host__ __device__ inline float pf_to_elev(uint64_t index, float min_elev) {
    return (float)index / 1000.0f + min_elev;
}
// A cell holding value v expands its neighbours in the reference's bucket
// flood only if v survives the bucket round trip.
__host__ __device__ inline bool pf_live(float v, float min_elev) {
    return pf_to_elev(pf_to_index(v, min_elev), min_elev) == v;
}

// Grid-wide barrier of the persistent kernels.  All workgroups must be
// co-resident: the grids are capped at the device's occupancy limit for the
// kernel, so every workgroup runs at once when the device is free of other
// work (other work can only delay a workgroup's start, and the barrier then
// waits for it).
//
// Cost model on a multi-XCD part: each XCD has its own L2, so making data
// written before the barrier visible to every workgroup after it takes an
// agent-scope release (L2 write-back) and acquire (L2 + L1 invalidate).
// Those are done once per workgroup, by thread 0, around a relaxed arrival
// counter and a relaxed poll of a generation word:
//  - __syncthreads() first: it waits for every wave's stores to reach L2,
//    so thread 0's release writes them back as well;
//  - thread 0's acquire invalidates the CU's L1, which all waves of the
//    workgroup share, and the closing __syncthreads() orders their loads
//    after it.
// Polling with acquire loads, or fencing in every thread, costs an L2
// invalidate per poll or per wave and made the barrier several times slower.
// Arrivals go to BAR_GROUPS counters (workgroup number modulo the group
// count), and each group's last arrival counts once at the top counter:
// with every workgroup arriving at once, 768 arrivals on one word took 18 us
// per barrier on an MI355X against 13 us spread over 16 words.  A group's
// last arrival resets its counter and releases before counting at the top,
// so the reset is seen before any arrival at the next barrier (those follow
// the generation bump, which follows the top count).
//
// The barrier also sums two counts over the workgroups (grid_sync_sum): the
// arrival words are 64-bit, arrivals in bits 0-15 and the two partial sums
// in bits 16-39 and 40-63, and the last arrival publishes the totals with
// the new generation in one 64-bit word (generation in bits 40-63, the sums
// in 20 bits each), so a waiting workgroup has them with the poll that ends
// its wait instead of one more dependent read after it.  The sums must stay
// below 2^20 (callers with more blocks than that read their counts after
// the barrier instead).
//
// The wait is bounded.  If co-residency is lost (another launch holds part
// of the device for long, or workgroups of a persistent launch on another
// stream of the same device take the places of some of this launch's), or
// the host raises the context's abort word, the waiting workgroup gives up:
// it raises the barrier block's abort word, records CAGEO_ERR_ABORTED in the
// launch's sticky status word and returns false.  Every other workgroup sees
// the abort word in its own wait (or at its next barrier) and returns false
// too, so the whole grid leaves its loops and ends normally.  A later
// persistent launch on the same barrier block ends at once while the abort
// word is up (grid_aborted); the host clears the word and the counters
// before its next call (k_reset_launch_state).
constexpr int BAR_STRIDE = 32;  // words: one 128-byte line per word
constexpr int BAR_GROUPS = 16;
constexpr int BAR_COUNT = 0;
constexpr int BAR_GEN = BAR_STRIDE;
constexpr int BAR_ABORT = 2 * BAR_STRIDE;
constexpr int BAR_GROUP = 3 * BAR_STRIDE;
constexpr int BAR_WORDS = (3 + BAR_GROUPS) * BAR_STRIDE;
// Polls between two checks of the device abort word and the clock, and
// between two reads of the host's abort word.  The host word lives in host
// memory: a read crosses PCIe, and one from every workgroup at every
// barrier cost the pit fill 2 ms of its 7 ms at 4096^2 on an MI355X.  So a
// launch reads it once at its start (workgroup 0, grid_aborted) and then
// only in waits that have lasted this many polls (a stalled launch).
constexpr uint32_t BAR_CHECK_EVERY = 64;
constexpr uint32_t BAR_HOST_EVERY = 16 * BAR_CHECK_EVERY;
// Default bound of one barrier wait, in seconds.
constexpr double BAR_DEFAULT_TIMEOUT_S = 5.0;

// What a persistent launch needs for its barriers.
struct GridBar {
    uint32_t* bar;                // BAR_WORDS device words
    const uint32_t* host_abort;   // host-mapped word raised by the host (may be null)
    uint64_t limit;               // bound of one wait, in wall_clock64() ticks
    uint32_t* status;             // the launch's sticky status word
};

__device__ inline bool bar_should_stop(const GridBar& gb, uint64_t t0, bool host) {
    if (__hip_atomic_load(&gb.bar[BAR_ABORT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        return true;
    if (wall_clock64() - t0 >= gb.limit) return true;
    return host && gb.host_abort &&
           __hip_atomic_load(gb.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
}

// Returns false when the launch is being abandoned (then every workgroup
// must leave its loops and end without another barrier).  The barrier is
// over workgroups 0 .. n-1 (all of the grid by default; a launch may go on
// with its first n workgroups once the others have ended).  add0 / add1
// (thread 0's) are summed over the workgroups into sum0 / sum1 (every
// thread's).
constexpr uint64_t BAR_SUM_LIMIT = 1u << 20;
__device__ inline bool grid_sync_sum(const GridBar& gb, uint32_t n, uint32_t add0, uint32_t add1,
                                     uint32_t& sum0, uint32_t& sum1) {
    __shared__ uint32_t s_sum[2];
    if (n == 0) n = gridDim.x;
    __syncthreads();
    int fail = 0;
    if (threadIdx.x == 0) {
        uint32_t* bar = gb.bar;
        uint64_t* gen = reinterpret_cast<uint64_t*>(&bar[BAR_GEN]);
        const uint64_t g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 40;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t groups = n < (uint32_t)BAR_GROUPS ? n : (uint32_t)BAR_GROUPS;
        const uint32_t grp = blockIdx.x % groups;
        const uint32_t members = (n - grp + groups - 1) / groups;
        uint64_t* gc = reinterpret_cast<uint64_t*>(&bar[BAR_GROUP + grp * BAR_STRIDE]);
        const uint64_t mine = 1u | ((uint64_t)add0 << 16) | ((uint64_t)add1 << 40);
        const uint64_t old = __hip_atomic_fetch_add(gc, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((old & 0xffffu) == members - 1) {
            const uint64_t group = old + mine;  // the group's arrivals and sums
            __hip_atomic_store(gc, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint64_t* top = reinterpret_cast<uint64_t*>(&bar[BAR_COUNT]);
            const uint64_t gm = 1u | (group & ~(uint64_t)0xffffu);
            const uint64_t told = __hip_atomic_fetch_add(top, gm, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((told & 0xffffu) == groups - 1) {
                // last arrival: reset the top counter, then open the next
                // generation with the totals
                const uint64_t all = told + gm;
                const uint64_t s0 = (all >> 16) & 0xffffffu, s1 = all >> 40;
                __hip_atomic_store(top, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(gen,
                                   (((g + 1) & 0xffffffu) << 40) | ((s0 & 0xfffffu) << 20) |
                                       (s1 & 0xfffffu),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const uint64_t t0 = wall_clock64();
        uint64_t now = 0;
        for (uint32_t it = 1;
             ((now = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 40) == g;
             ++it) {
            // a zero bound gives up at the first poll that finds the
            // barrier closed
            if ((it % BAR_CHECK_EVERY == 0 || gb.limit == 0) &&
                bar_should_stop(gb, t0, it % BAR_HOST_EVERY == 0)) {
                __hip_atomic_store(&bar[BAR_ABORT], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                atomicCAS(gb.status, 0u, (uint32_t)(-CAGEO_ERR_ABORTED));
                fail = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (!fail) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_sum[0] = (uint32_t)((now >> 20) & 0xfffffu);
        s_sum[1] = (uint32_t)(now & 0xfffffu);
    }
    const bool ok = !__syncthreads_or(fail);
    sum0 = s_sum[0];
    sum1 = s_sum[1];
    return ok;
}
__device__ inline bool grid_sync(const GridBar& gb, uint32_t n = 0) {
    uint32_t s0, s1;
    return grid_sync_sum(gb, n, 0u, 0u, s0, s1);
}

// True (in every thread of the workgroup) when an earlier launch on this
// barrier block was abandoned and the host has not cleared it yet, or (read
// by workgroup 0) the host's abort word is raised: the launch must end at
// once.  Workgroups already past this point see the device word in their
// next barrier wait.
__device__ inline bool grid_aborted(const GridBar& gb) {
    bool stop = false;
    if (threadIdx.x == 0) {
        stop = __hip_atomic_load(&gb.bar[BAR_ABORT], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
        if (!stop && blockIdx.x == 0 && gb.host_abort &&
            __hip_atomic_load(gb.host_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u) {
            __hip_atomic_store(&gb.bar[BAR_ABORT], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            atomicCAS(gb.status, 0u, (uint32_t)(-CAGEO_ERR_ABORTED));
            stop = true;
        }
    }
    return __syncthreads_or(stop);
}

// Words of the control block (d_ctrl), shared by the persistent kernels.
enum : int {
    CTL_COUNT = 0,      // 3 worklist counts
    CTL_CERT = 3,       // 3 rotating per-round fixed-cell counters (pit fill)
    CTL_STATUS = 6,     // sticky error: -CAGEO_ERR_* of the first failing stage
    CTL_EPOCH = 7,      // epoch counter of the pit-fill worklists
    CTL_ROUNDS = 8,
    CTL_PASSES = 9,
    CTL_FLAG = 10,      // 3 words, free (pass-end flags travel with the barrier)
    CTL_EPOCH_FR = 13,  // free (flat-resolution epochs start over each launch)
    CTL_FR_PASSES = 14, // flat-resolution passes of the last launch
    CTL_TR_ROUNDS = 15, // exit-jump rounds of the last tracing launch
    CTL_PHASE = 16,     // 3 phase-list counts (pit fill)
    CTL_CLAIM = 19,     // 3 rotating tile-claim counters of the passes
    CTL_COMPACT = 22,   // count of the pit fill's compacted fill list
    CTL_COUNTB = 23,    // 3 back counts of the pit fill's two-ended pass lists
    CTL_PHASEB = 26,    // 3 back counts of its phase lists
    CTL_COMPACTB = 29,  // back count of its compacted fill list
    CTL_RESIDENT = 30,  // pit fill: open tiles handed to the resident rounds (0: none)
    CTL_RROUND = 31,    // pit fill: the round the resident rounds go on from
    CTL_FR_DONE = 32,   // flow routing: 1 once the flat passes of the launch converged
    CTL_WORDS = 33
};

// Several rasters in one persistent launch: raster r owns the global 64x64
// block numbers [base[r], base[r+1]), row-major within the raster;
// bmap[g] = (raster, block number within it).
//
// Sweeps over every cell of every raster go by units of UNIT_ROWS rows x 64
// columns of a block, one wave per unit, grid-strided: even shares whatever
// the raster sizes, and the raster lookup is paid once per unit.
constexpr int UNIT_ROWS = 4;
constexpr int UNITS_PER_BLOCK = 64 / UNIT_ROWS;
__device__ inline long long wave_unit_first() {
    return (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
}
__device__ inline long long wave_unit_stride() { return (long long)gridDim.x * (blockDim.x / 64); }

// Device state of the multi-raster persistent launches of a batch (the
// per-raster buffers stay in each raster's context).
struct MultiState {
    int nr = 0, total = 0;            // rasters, global 64x64 blocks
    int max_cells = 0;                // largest raster (cells)
    int* d_base = nullptr;            // nr + 1 block offsets
    int2* d_bmap = nullptr;           // per block: (raster, block within it)
    uint2* d_rinfo = nullptr;         // per raster: pit-fill minimum, domain ok
    int* d_work = nullptr;            // 7 lists of `total` entries
    uint32_t* d_mark_pf = nullptr;    // per-block epoch stamps (pit fill)
    uint32_t* d_mark_fr = nullptr;    // per-block epoch stamps (flats)
    uint32_t* d_rows = nullptr;       // per block: dirty-row masks, 2 pass parities x 64 bits
    uint32_t* d_topen = nullptr;      // per block: holds OPEN pit-fill cells
    uint32_t* d_tcost = nullptr;      // per block: row steps of its last pit-fill solve
    uint32_t* d_ctrl = nullptr;       // CTL_WORDS
    uint32_t* d_cert = nullptr;       // 3 rotating counters per raster
    uint32_t* d_bar = nullptr;        // grid barrier words
    const uint32_t* d_habort = nullptr;  // host abort word (the batch's)
    uint64_t bar_limit = 0;
    int pf_cap = 0, fr_cap = 0, tr_cap = 0;  // co-resident workgroups of the all-tile kernels
    void* d_pf = nullptr;             // raster tables of the three kernels
    void* d_fr = nullptr;
    void* d_tr = nullptr;
    int pf_grid = 0, fr_grid = 0, tr_grid = 0;
};

// Worklist helpers --------------------------------------------------------
int read_scalars(Context* c);
// Barrier of a launch on the context's (or the batch's) device state.
inline GridBar context_gridbar(Context* c) {
    return GridBar{c->d_bar, c->d_habort, c->bar_limit, c->d_ctrl + CTL_STATUS};
}
inline GridBar multi_gridbar(MultiState* m) {
    return GridBar{m->d_bar, m->d_habort, m->bar_limit, m->d_ctrl + CTL_STATUS};
}
// Ticks of wall_clock64() in `seconds` on `device`.
int barrier_ticks(int device, double seconds, uint64_t* ticks);
// Clears a launch state after any earlier launch: the sticky status word,
// the barrier counters and the barrier abort word (one small kernel, so
// that it may be captured into a graph).
int reset_launch_state(hipStream_t s, uint32_t* status, uint32_t* bar, uint32_t* scal = nullptr);

// Stage drivers implemented in the stage translation units -----------------
int launch_random_field(Context* c, float* out, uint64_t seed, uint32_t realization,
                        uint32_t stream_id, float mean, float std_dev,
                        const uint32_t* realization_base = nullptr);
int launch_convolve(Context* c, float* field, float* buf);
// Convolution whose result may stay in `buf` (exponential model): returns in
// *result the buffer that holds it, saving the copy back into `field`.
int launch_convolve_to(Context* c, float* field, float* buf, float** result);
int launch_convolve_fast(Context* c, const float* field, float* out);
int lr_cols_geometry(int r, int* LC, size_t* lds);
int lr_rows_geometry(int r, int* LP, int* S, size_t* lds);
int prepare_low_rank(Context* c);
bool fft_supported(int W, int H, int r);
int prepare_fft(Context* c);
void release_fft(Context* c);
int launch_convolve_fft(Context* c, const float* field, float* out);
constexpr size_t LDS_LIMIT = 160 * 1024;
int conv_2d_config(int r, int* ty, int* S, size_t* lds_bytes);
int launch_realistic(Context* c, const float* dem, float* field, float nodata);
int launch_burn(Context* c, const float* zin, float* zout, const uint8_t* sflags,
                float nodata_dem);
int prepare_offsets(Context* c, int radius);
int launch_stream_flags(Context* c, const float* stream, float nodata_stream, uint8_t* sflags,
                        int32_t* slabel);
int launch_realistic_burn_quant(Context* c, const float* dem, const float* e, float* zq);
size_t burn_blocks(int W, int H);
int pit_fill_quantized(Context* c, float* zq, float* w, float nodata, bool stats_ready);
int check_status(Context* c);
int reset_minmax(Context* c);
// One realization from the error field in c->d_a (stages a3..a9); with
// `evs`, events evs[2..7] are recorded after each stage.
int realization_body(Context* c, float* mc, hipEvent_t* evs);
// Its first part: stages a3..a5 and the quantisation into c->d_z.
int realization_pre(Context* c, hipEvent_t* evs);
int pit_fill_grid_cap(Context* c, int* cap);
int pit_fill_resident_cap(Context* c, int* cap);
int flow_routing_grid_cap(Context* c, int* cap);
int flow_tracing_grid_cap(Context* c, int* cap);
int clear_status(Context* c);
int launch_quantize(Context* c, float* dem, float nodata);
int flow_routing(Context* c, const float* z, int8_t* fd);
int flow_tracing(Context* c, const int8_t* fd, const uint8_t* sflags, const int32_t* slabel,
                 int32_t* lab, int64_t* long_paths);
int launch_border(Context* c, float* mc, const int32_t* lab);
int launch_to_bytes(Context* c, uint8_t* out, const float* in);
// Multi-raster launches over the contexts of a batch (tables built once by
// multi_setup; each context's working buffers are used as in the one-raster
// stages: pit fill d_z -> d_w, routing d_w -> d_fd, tracing d_fd -> d_lab).
int multi_setup(MultiState* m, Context* const* ctxs, int nr);
void multi_release(MultiState* m);
int pit_fill_multi(MultiState* m, hipStream_t s);
int flow_routing_multi(MultiState* m, hipStream_t s);
int flow_tracing_multi(MultiState* m, hipStream_t s);
int pit_fill_tables(MultiState* m, Context* const* ctxs);
int flow_routing_tables(MultiState* m, Context* const* ctxs);
int flow_tracing_tables(MultiState* m, Context* const* ctxs);

}  // namespace cageo

struct cageo_context {
    cageo::Context c;
};
