// Flow tracing: FlowTracingAlgorithm<T,U,V>::execute, FlowTracingAlgorithm.h:
// 52-69; parity target FlowTracingAlgorithmCPUOpenMP.h:34-134.
//
// The reference walks every untraced cell along its directions until it
// leaves the grid (OUTSIDE), steps onto a stream cell (label = -stream as
// int) or onto an already labelled cell (copies it); stream cells themselves
// are LAKE; a walk longer than 100 001 steps (a cell with no direction, or a
// cycle) leaves its path at 0.  A cell's label is therefore the label of its
// path's terminal, computed here by pointer jumping in two levels:
//   1. per 64x64 tile, in LDS: every cell's pointer jumps along the in-tile
//      part of its path (12 doubling rounds cover 4096 cells) and ends as a
//      terminal label, "stuck" (its path cycles inside the tile, e.g. a cell
//      without a direction), or an exit: the first cell outside the tile;
//   2. over the grid: exits jump to the entries of their exit cells
//      (ping-pong buffers) until no exit is left; a chain still moving after
//      ceil(log2 n)+1 rounds is a cycle through several tiles (stuck).
// Entries are 64 bits: kind (2 bits) | steps (30 bits) | value (32 bits), so
// the step count to the terminal is known and cells whose terminal is more
// than 100 001 steps away (where the reference's result depends on its
// raster order) are counted.  performConnectFlowPath (:115-134) then maps a
// label >= 0 to label[label] when that is negative.
#include "common.hpp"

namespace cageo {

namespace {

constexpr uint32_t K_LOCAL = 0, K_RESOLVED = 1, K_EXIT = 2, K_STUCK = 3;
constexpr uint64_t STEP_MASK = (1ull << 30) - 1;

__device__ inline uint64_t pack(uint32_t kind, uint32_t steps, int32_t value) {
    return ((uint64_t)kind << 62) | ((uint64_t)(steps & STEP_MASK) << 32) |
           (uint64_t)(uint32_t)value;
}
__device__ inline uint32_t kind_of(uint64_t e) { return (uint32_t)(e >> 62); }
__device__ inline uint32_t steps_of(uint64_t e) { return (uint32_t)((e >> 32) & STEP_MASK); }
__device__ inline int32_t value_of(uint64_t e) { return (int32_t)(uint32_t)e; }

// Follow entry e one hop through entry f (the entry of e's target).
__device__ inline uint64_t hop(uint64_t e, uint64_t f) {
    uint32_t s = steps_of(e) + steps_of(f);
    if (s > STEP_MASK) s = STEP_MASK;
    const uint32_t k = kind_of(f);
    if (k == K_STUCK) return pack(K_STUCK, 0, 0);
    return pack(k, s, value_of(f));
}

__device__ void trace_tile(uint64_t* L, const int8_t* __restrict__ fd,
                           const uint8_t* __restrict__ sf, const int32_t* __restrict__ slabel,
                           uint64_t* __restrict__ ent, int W, int H, int ntx, int t) {
    const int x0 = (t % ntx) * TILE, y0 = (t / ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = x0 + lx;

    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const int gy = y0 + ly;
        uint64_t e = pack(K_STUCK, 0, 0);
        if (gx < W && gy < H) {
            const size_t i = (size_t)gy * W + gx;
            if (sf[i] & SF_TRACE) {
                e = pack(K_RESOLVED, 0, CAGEO_FLOWTRACE_LAKE);
            } else {
                const char2 d = reinterpret_cast<const char2*>(fd)[i];
                const int nx = gx + d.x, ny = gy + d.y;
                if (nx < 0 || ny < 0 || nx >= W || ny >= H) {
                    e = pack(K_RESOLVED, 1, CAGEO_FLOWTRACE_OUTSIDE);
                } else {
                    const size_t ni = (size_t)ny * W + nx;
                    if (sf[ni] & SF_TRACE) {
                        e = pack(K_RESOLVED, 1, slabel[ni]);
                    } else if (d.x == 0 && d.y == 0) {
                        e = pack(K_STUCK, 0, 0);
                    } else if (nx >= x0 && ny >= y0 && nx < x0 + TILE && ny < y0 + TILE) {
                        e = pack(K_LOCAL, 1, (ny - y0) * TILE + (nx - x0));
                        local |= 1u << k;
                    } else {
                        e = pack(K_EXIT, 1, (int32_t)ni);
                    }
                }
            }
        }
        L[ly * TILE + lx] = e;
    }
    __syncthreads();

    // Jacobi doubling rounds: read all targets, barrier, write.
    for (int round = 0; round < 13; ++round) {
        if (!__syncthreads_or(local != 0)) break;
        uint64_t nv[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            const int li = (strip * STRIP + k) * TILE + lx;
            nv[k] = 0;
            if (local & (1u << k)) {
                const uint64_t e = L[li];
                nv[k] = hop(e, L[value_of(e)]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            if (local & (1u << k)) {
                const int li = (strip * STRIP + k) * TILE + lx;
                L[li] = nv[k];
                if (kind_of(nv[k]) != K_LOCAL) local &= ~(1u << k);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const int gy = y0 + ly;
        if (gx < W && gy < H) {
            uint64_t e = L[ly * TILE + lx];
            if (kind_of(e) == K_LOCAL) e = pack(K_STUCK, 0, 0);  // cycle inside the tile
            ent[(size_t)gy * W + gx] = e;
        }
    }
    __syncthreads();
}


__device__ inline uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct TrArgs {
    const int8_t* fd;
    const uint8_t* sf;
    const int32_t* slabel;
    uint64_t* ent[2];
    int32_t* lab;
    int W, H, ntx, ntiles;
    int max_rounds;
    uint32_t* ctrl;
    uint32_t* bar;  // grid barrier words
    uint32_t* scal;
};

// Tile pass, exit jumping, labels and the connect step in one persistent
// launch.  Round r of the exit jumping raises flag r % 3 while exits remain
// and clears flag (r+1) % 3, whose last read (end of round r-2) is separated
// from the clear by barriers.
__global__ __launch_bounds__(256) void k_tr_persistent(TrArgs a) {
    __shared__ uint64_t L[TILE * TILE];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl[CTL_FLAG + 0] = 0;
        a.ctrl[CTL_FLAG + 1] = 0;
        a.ctrl[CTL_FLAG + 2] = 0;
        a.scal[SC_LONGPATH] = 0;
    }
    for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x)
        trace_tile(L, a.fd, a.sf, a.slabel, a.ent[0], a.W, a.H, a.ntx, t);
    grid_sync(a.bar);

    const size_t n = (size_t)a.W * a.H;
    int cur = 0;
    int rounds = 0;
    for (int r = 0; r < a.max_rounds; ++r) {
        rounds = r + 1;
        const uint64_t* src = a.ent[cur];
        uint64_t* dst = a.ent[cur ^ 1];
        uint32_t pending = 0;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (size_t)gridDim.x * blockDim.x) {
            uint64_t e = src[i];
            if (kind_of(e) == K_EXIT) {
                e = hop(e, src[value_of(e)]);
                pending |= kind_of(e) == K_EXIT;
            }
            dst[i] = e;
        }
        if (__any(pending) && (threadIdx.x & 63) == 0) atomicOr(&a.ctrl[CTL_FLAG + r % 3], 1u);
        if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl[CTL_FLAG + (r + 1) % 3] = 0;
        grid_sync(a.bar);
        cur ^= 1;
        if (!ld(&a.ctrl[CTL_FLAG + r % 3])) break;
    }

    uint32_t longp = 0;
    const uint64_t* fin = a.ent[cur];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t e = fin[i];
        int32_t v = 0;
        if (kind_of(e) == K_RESOLVED) {
            v = value_of(e);
            if (steps_of(e) > 100001u) ++longp;
        }
        a.lab[i] = v;
    }
    for (int off = 32; off > 0; off >>= 1) longp += (uint32_t)__shfl_xor((int)longp, off);
    if ((threadIdx.x & 63) == 0 && longp) atomicAdd(&a.scal[SC_LONGPATH], longp);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl[CTL_TR_ROUNDS] = (uint32_t)rounds;  // host only
    grid_sync(a.bar);
    // performConnectFlowPath: label[0] is final and never changes here.
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int32_t v = a.lab[i];
        if (v >= 0 && (size_t)v < n) {
            const int32_t t = a.lab[v];
            if (t < 0) a.lab[i] = t;
        }
    }
}

}  // namespace

// Needs two 64-bit rasters: c->d_ptr and c->d_ptr2.  The long-path count is
// left in the scalar block.
// Device function of the persistent kernel (to find its node in a graph).
const void* flow_tracing_kernel() { return reinterpret_cast<const void*>(k_tr_persistent); }

// Co-resident workgroups of the persistent kernel on this device.
int flow_tracing_grid_cap(Context* c, int* cap) {
    int per_cu = 0;
    CAGEO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_tr_persistent), TILE_THREADS, 0));
    hipDeviceProp_t prop;
    CAGEO_HIP_TRY(hipGetDeviceProperties(&prop, c->device));
    if (per_cu < 1) return CAGEO_ERR_HIP;
    *cap = per_cu * prop.multiProcessorCount;
    return CAGEO_OK;
}

int flow_tracing(Context* c, const int8_t* fd, const uint8_t* sflags, const int32_t* slabel,
                 int32_t* lab, int64_t* long_paths) {
    if (c->tr_grid <= 0) CAGEO_TRY(flow_tracing_grid_cap(c, &c->tr_grid));
    TrArgs a;
    a.fd = fd;
    a.sf = sflags;
    a.slabel = slabel;
    a.ent[0] = c->d_ptr;
    a.ent[1] = c->d_ptr2;
    a.lab = lab;
    a.W = c->W;
    a.H = c->H;
    a.ntx = c->ntx;
    a.ntiles = c->ntiles;
    int max_rounds = 2;
    for (size_t m = 1; m < c->n; m <<= 1) ++max_rounds;
    a.max_rounds = max_rounds;
    a.ctrl = c->d_ctrl;
    a.bar = c->d_bar;
    a.scal = c->d_scal;
    // ordinary launch within the occupancy cap (see pit_fill_quantized)
    hipLaunchKernelGGL(k_tr_persistent, dim3(c->tr_grid), dim3(TILE_THREADS), 0, c->stream, a);
    CAGEO_HIP_TRY(hipGetLastError());
    if (long_paths) {
        CAGEO_TRY(read_scalars(c));
        *long_paths = (int64_t)c->h_scal[SC_LONGPATH];
    }
    return CAGEO_OK;
}

}  // namespace cageo
