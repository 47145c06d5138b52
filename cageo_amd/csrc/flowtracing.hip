// Flow tracing: FlowTracingAlgorithm<T,U,V>::execute, FlowTracingAlgorithm.h:
// 52-69; parity target FlowTracingAlgorithmCPUOpenMP.h:34-134.
//
// The reference walks every untraced cell along its directions until it
// leaves the grid (OUTSIDE), steps onto a stream cell (label = -stream as
// int) or onto an already labelled cell (copies it); stream cells themselves
// are LAKE; a walk longer than 100 001 steps (a cell with no direction, or a
// cycle) leaves its path at 0.  A cell's label is therefore the label of its
// path's terminal, which pointer jumping computes in O(log path) passes:
//   entry(i) = terminal label    if i's next cell decides it, else
//              pointer to next cell;  entry(i) <- entry(entry(i)).
// Entries are packed in 64 bits (flags | steps | value) so one atomic load
// reads a consistent pointer/step pair while other lanes update in place.
// A pointer that stops advancing points into a cell without a direction
// (stuck, label 0); pointers still moving after ceil(log2 n)+1 passes are on
// a cycle (label 0).  Cells whose terminal is more than 100 001 steps away
// are counted: there the reference's result depends on its raster order.
// performConnectFlowPath (:115-134) then maps label >= 0 to label[label]
// when that is negative.
#include "common.hpp"

namespace cageo {

namespace {

constexpr uint64_t E_RESOLVED = 1ull << 63;
constexpr uint64_t E_STUCK = 1ull << 62;
constexpr int STEP_SHIFT = 32;
constexpr uint64_t STEP_MASK = (1ull << 30) - 1;

__device__ inline uint64_t pack(bool resolved, uint32_t steps, int32_t value) {
    return (resolved ? E_RESOLVED : 0ull) | ((uint64_t)(steps & STEP_MASK) << STEP_SHIFT) |
           (uint64_t)(uint32_t)value;
}
__device__ inline uint32_t steps_of(uint64_t e) { return (uint32_t)((e >> STEP_SHIFT) & STEP_MASK); }
__device__ inline int32_t value_of(uint64_t e) { return (int32_t)(uint32_t)e; }

__global__ __launch_bounds__(256) void k_trace_init(const int8_t* __restrict__ fd,
                                                    const uint8_t* __restrict__ sf,
                                                    const int32_t* __restrict__ slabel,
                                                    uint64_t* __restrict__ ent, int W, int H) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    uint64_t e;
    if (sf[i] & SF_TRACE) {
        e = pack(true, 0, CAGEO_FLOWTRACE_LAKE);
    } else {
        const char2 d = reinterpret_cast<const char2*>(fd)[i];
        const int nx = x + d.x, ny = y + d.y;
        if (nx < 0 || ny < 0 || nx >= W || ny >= H) {
            e = pack(true, 1, CAGEO_FLOWTRACE_OUTSIDE);
        } else {
            const size_t ni = (size_t)ny * W + nx;
            if (sf[ni] & SF_TRACE)
                e = pack(true, 1, slabel[ni]);
            else
                e = pack(false, 1, (int32_t)ni);
        }
    }
    ent[i] = e;
}

__global__ __launch_bounds__(256) void k_trace_jump(uint64_t* __restrict__ ent, int64_t n,
                                                    uint32_t* __restrict__ scal) {
    uint32_t changed = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = __hip_atomic_load(&ent[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e & E_RESOLVED) continue;
        const int32_t p = value_of(e);
        const uint64_t f = __hip_atomic_load(&ent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t ne;
        uint32_t s = steps_of(e) + steps_of(f);
        if (s > STEP_MASK) s = STEP_MASK;
        if (f & E_RESOLVED) {
            ne = (f & E_STUCK) ? (E_RESOLVED | E_STUCK) : pack(true, s, value_of(f));
        } else if (value_of(f) == p) {
            ne = E_RESOLVED | E_STUCK;  // next cell points at itself
        } else {
            ne = pack(false, s, value_of(f));
        }
        if (ne != e) {
            __hip_atomic_store(&ent[i], ne, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            changed = 1;
        }
    }
    if (changed) atomicOr(&scal[SC_CHANGED], 1u);
}

__global__ __launch_bounds__(256) void k_trace_final(const uint64_t* __restrict__ ent,
                                                     int32_t* __restrict__ lab, int64_t n,
                                                     uint32_t* __restrict__ scal) {
    uint32_t longp = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t e = ent[i];
        int32_t v = 0;
        if ((e & E_RESOLVED) && !(e & E_STUCK)) {
            v = value_of(e);
            if (steps_of(e) > 100001u) ++longp;
        }
        lab[i] = v;
    }
    if (longp) atomicAdd(&scal[SC_LONGPATH], longp);
}

__global__ __launch_bounds__(256) void k_trace_connect(int32_t* __restrict__ lab, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = lab[i];
        if (v >= 0 && (int64_t)v < n) {
            const int32_t t = lab[v];
            if (t < 0) lab[i] = t;
        }
    }
}

inline unsigned grid_1d(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

}  // namespace

int flow_tracing(Context* c, const int8_t* fd, const uint8_t* sflags, const int32_t* slabel,
                 int32_t* lab, int64_t* long_paths) {
    const int W = c->W, H = c->H;
    hipLaunchKernelGGL(k_trace_init, dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, c->stream,
                       fd, sflags, slabel, c->d_ptr, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    int max_rounds = 2;
    for (size_t m = 1; m < c->n; m <<= 1) ++max_rounds;
    for (int r = 0; r < max_rounds; ++r) {
        CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_CHANGED, 0, sizeof(uint32_t), c->stream));
        hipLaunchKernelGGL(k_trace_jump, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, c->d_ptr,
                           (int64_t)c->n, c->d_scal);
        CAGEO_HIP_TRY(hipGetLastError());
        CAGEO_TRY(read_scalars(c));
        if (!c->h_scal[SC_CHANGED]) break;
    }
    CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_LONGPATH, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_trace_final, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, c->d_ptr,
                       lab, (int64_t)c->n, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_trace_connect, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, lab,
                       (int64_t)c->n);
    CAGEO_HIP_TRY(hipGetLastError());
    if (long_paths) {
        CAGEO_TRY(read_scalars(c));
        *long_paths = (int64_t)c->h_scal[SC_LONGPATH];
    }
    return CAGEO_OK;
}

}  // namespace cageo
