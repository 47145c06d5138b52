// Surface stages that turn the error field into the DEM the pit fill sees:
//  - RealisticDemAlgorithm<T,U>::execute, RealisticDemAlgorithm.h:49-70
//    (AddSurfacesAlgorithmCPUOpenMP.h:36-56 then
//    RemoveNegativeElevationsAlgorithmCPUOpenMP.h:34-48);
//  - StreamBurningAlgorithm<T,U>::execute, StreamBurningAlgorithm.h:116-134
//    with performStreamBurning, StreamBurningAlgorithmCPUOpenMP.h:32-82;
//  - the mm quantisation that opens performPitFilling,
//    PitFillingAlgorithmCPUOpenMP.h:84-106.
//
// Burning semantics.  The reference updates z in place while scanning each
// cell's circular search array; with one thread in raster order a stream
// neighbour n that precedes the cell has already been set to 0, one that
// follows still holds its realistic-DEM value.  Non-stream cell i takes
//   z_i * factor(o)  for the first offset o (order of
//                    StreamBurningAlgorithm.h:140-167) whose cell n is in the
//                    grid, is a stream cell, and has seen(n) != nodata_dem,
//                    seen(n) = n < i ? 0 : z_n;
//   z_i              otherwise;
// stream cells become 0.  Evaluated out of place here so every cell reads the
// unburned values.
#include <cmath>
#include <list>
#include <vector>

#include "tiles.hpp"

namespace cageo {

namespace {


__global__ __launch_bounds__(256) void k_stream_flags(const float* __restrict__ stream,
                                                      int64_t n, float nodata_stream,
                                                      uint8_t* __restrict__ sflags,
                                                      int32_t* __restrict__ slabel) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s = stream[i];
        uint8_t f = 0;
        if (s != nodata_stream) f |= SF_BURN;  // StreamBurningAlgorithmCPUOpenMP.h:46
        if (s >= 0.0f) f |= SF_TRACE;          // FlowTracingAlgorithmCPUOpenMP.h:52,88
        sflags[i] = f;
        // FlowTracingAlgorithmCPUOpenMP.h:89: label = -stream converted to int
        slabel[i] = (s >= 0.0f) ? (int32_t)(-s) : 0;
    }
}

__device__ inline float realistic_value(float d, float e, float nodata) {
    float v = (d != nodata) ? d + e : e;
    if (v <= 0.0f && v != nodata) v = 1.0f;
    return v;
}

__global__ __launch_bounds__(256) void k_realistic(const float* __restrict__ dem,
                                                   float* __restrict__ field, int64_t n,
                                                   float nodata) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        field[i] = realistic_value(dem[i], field[i], nodata);
}

// Burning works on LDS tiles of 64 x 16 cells with an R-cell halo holding
// the unburned values and, per cell, the two facts the search asks about a
// neighbour n, given that n is in the grid and a stream cell:
//   BF_BURN  n is a stream cell;
//   BF_LATE  n is a stream cell whose unburned value is not nodata_dem.
// A neighbour that precedes the cell in raster order (offset dy < 0, or
// dy == 0 and dx < 0: for in-grid cells the reference's n < i reduces to
// the offset alone) has been burned to 0 already and qualifies when
// 0 != nodata_dem, i.e. by BF_BURN under that launch constant; a later one
// qualifies by BF_LATE.  Halo cells outside the grid hold no flag, which
// is the reference's in-grid test.  Each offset's LDS step and the flag it
// needs are packed into one word of a table staged in LDS.
constexpr int BT_X = 64, BT_Y = 16;
constexpr int BH_MAX = CAGEO_MAX_BEVEL_RADIUS;
constexpr int BL_W = BT_X + 2 * BH_MAX, BL_H = BT_Y + 2 * BH_MAX;
constexpr int BOFF_MAX = (2 * BH_MAX + 1) * (2 * BH_MAX + 1);
constexpr uint8_t BF_BURN = 1, BF_LATE = 2;

// REAL: the tile's values are realistic_value(z0, z1) (the fused Monte
// Carlo pass), else z0 itself.
template <bool REAL>
__device__ inline void burn_tile(int W, int H, const uint8_t* __restrict__ sf,
                                 const int* __restrict__ off, const float* __restrict__ fac,
                                 int n_off, int R, float nodata_dem,
                                 const float* __restrict__ z0, const float* __restrict__ z1,
                                 float out[4], int gx_out[4], int gy_out[4]) {
    __shared__ float zt[BL_W * BL_H];
    __shared__ uint8_t ft[BL_W * BL_H];
    __shared__ int s_step[BOFF_MAX];  // LDS step << 2 | flag the offset needs
    __shared__ float s_fac[BOFF_MAX];
    constexpr int ITER = (BL_W * BL_H + 255) / 256;  // loads per thread at the largest radius
    // Blocks go round robin over the 8 XCDs; renumbering them so that each
    // XCD takes a contiguous band of tiles keeps the neighbours whose cache
    // lines hold this tile's halo in the same L2 (a bijection on the block
    // numbers when their count is a multiple of 8).
    const unsigned nb = gridDim.x * gridDim.y;
    unsigned t = blockIdx.y * gridDim.x + blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int x0 = (int)(t % gridDim.x) * BT_X, y0 = (int)(t / gridDim.x) * BT_Y;
    const int lw = BT_X + 2 * R, lh = BT_Y + 2 * R;
    // All of a thread's loads are issued before any of them is used: a cell
    // outside the grid or the halo of radius R loads cell 0 (cached) and is
    // replaced afterwards, so no load sits in a branch that would have to
    // wait for it.
    float av[ITER], bv[ITER];
    uint8_t fv[ITER];
    bool in[ITER];
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int k = threadIdx.x + it * 256;
        const int ly = k / lw, lx = k - ly * lw;
        const int gx = x0 - R + lx, gy = y0 - R + ly;
        in[it] = k < lw * lh && gx >= 0 && gy >= 0 && gx < W && gy < H;
        const size_t gi = in[it] ? (size_t)gy * W + gx : 0;
        av[it] = z0[gi];
        bv[it] = REAL ? z1[gi] : 0.0f;
        fv[it] = sf[gi];
    }
    if ((int)threadIdx.x < n_off) {
        const int o = threadIdx.x, ox = off[2 * o], oy = off[2 * o + 1];
        const bool early = oy < 0 || (oy == 0 && ox < 0);
        const int need = early ? (nodata_dem != 0.0f ? BF_BURN : 0) : BF_LATE;
        s_step[o] = (oy * BL_W + ox) * 4 + need;
        s_fac[o] = fac[o];
    }
#pragma unroll
    for (int it = 0; it < ITER; ++it) {
        const int k = threadIdx.x + it * 256;
        if (k < lw * lh) {
            const int ly = k / lw, lx = k - ly * lw;
            float v = 0.0f;
            uint8_t f = 0;
            if (in[it]) {
                v = REAL ? realistic_value(av[it], bv[it], nodata_dem) : av[it];
                if (fv[it] & SF_BURN) f = BF_BURN | (v != nodata_dem ? BF_LATE : 0);
            }
            zt[ly * BL_W + lx] = v;
            ft[ly * BL_W + lx] = f;
        }
    }
    __syncthreads();
    // Each thread's four cells (one column, four rows) search side by side:
    // per offset one table read and four independent flag reads.  A cell
    // takes the first offset, in the search array's order, whose cell
    // qualifies; stream cells become 0.
    const int lx = threadIdx.x & 63;
    int c[4];
    bool done[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ly = (threadIdx.x >> 6) * 4 + j;
        gx_out[j] = x0 + lx;
        gy_out[j] = y0 + ly;
        c[j] = (ly + R) * BL_W + lx + R;
        out[j] = zt[c[j]];
        done[j] = (ft[c[j]] & BF_BURN) != 0;
        if (done[j]) out[j] = 0.0f;
    }
    for (int o = 0; o < n_off; ++o) {
        if (!__any(!(done[0] && done[1] && done[2] && done[3]))) break;
        const int s = s_step[o];
        const float f = s_fac[o];
        uint8_t q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = ft[c[j] + (s >> 2)];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!done[j] && (q[j] & s & 3)) {
                out[j] = out[j] * f;
                done[j] = true;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_burn(const float* __restrict__ zin,
                                              float* __restrict__ zout,
                                              const uint8_t* __restrict__ sf,
                                              const int* __restrict__ off,
                                              const float* __restrict__ fac, int n_off, int R,
                                              int W, int H, float nodata_dem) {
    float v[4];
    int gx[4], gy[4];
    burn_tile<false>(W, H, sf, off, fac, n_off, R, nodata_dem, zin, nullptr, v, gx, gy);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (gx[j] < W && gy[j] < H) zout[(size_t)gy[j] * W + gx[j]] = v[j];
}

// Quantisation of PitFillingAlgorithmCPUOpenMP.h:84-106 for one value.
__device__ inline float quantize_mm(float v, float nodata, uint32_t* bad) {
    if (v == nodata) return v;
    if (v < 0.0f || v != v) {
        *bad = 1u;
        return v;
    }
    uint64_t q = pf_to_index(v, 0.0f);
    if (v != 0.0f && q == 0) q = 1;
    return pf_to_elev(q, 0.0f);
}

// (min, max) of ordered uints: "self after prev" of the wave scan.
__device__ inline void minmax_after(int& a, int& b, int pa, int pb) {
    a = (int)min((uint32_t)a, (uint32_t)pa);
    b = (int)max((uint32_t)b, (uint32_t)pb);
}

// Block-wide min/max of ordered uints (wave scans on DPP, then the four
// waves' results through LDS), then at most one atomic per word per block:
// a block whose value cannot improve the global word (which only moves
// one way) skips it, so the atomics do not queue on the three words.
__device__ inline void block_minmax(uint32_t mn, uint32_t mx, uint32_t bad, uint32_t* scal) {
    __shared__ uint32_t s_mn[4], s_mx[4], s_bad[4];
    int a = (int)mn, b = (int)mx;
    CAGEO_SCAN_UP(dpp_i, a, b, (int)0xffffffffu, 0, minmax_after);  // lane 63 holds the wave's
    const int w = threadIdx.x >> 6;
    const uint32_t wb = __any(bad != 0u) ? 1u : 0u;
    if ((threadIdx.x & 63) == 63) {
        s_mn[w] = (uint32_t)a;
        s_mx[w] = (uint32_t)b;
        s_bad[w] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m0 = s_mn[0], m1 = s_mx[0], bb = s_bad[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            m0 = min(m0, s_mn[k]);
            m1 = max(m1, s_mx[k]);
            bb |= s_bad[k];
        }
        if (m0 < __hip_atomic_load(&scal[SC_MIN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&scal[SC_MIN], m0);
        if (m1 > __hip_atomic_load(&scal[SC_MAX], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&scal[SC_MAX], m1);
        if (bb) atomicOr(&scal[SC_BAD], bb);
    }
}

// The block's (min, max, bad) into part[0..2] (thread 0; plain stores, no
// wait on the global words): reduced by k_minmax_finish.
__device__ inline void block_minmax_store(uint32_t mn, uint32_t mx, uint32_t bad, uint32_t* part) {
    __shared__ uint32_t s_mn[4], s_mx[4], s_bad[4];
    int a = (int)mn, b = (int)mx;
    CAGEO_SCAN_UP(dpp_i, a, b, (int)0xffffffffu, 0, minmax_after);  // lane 63 holds the wave's
    const int w = threadIdx.x >> 6;
    const uint32_t wb = __any(bad != 0u) ? 1u : 0u;
    if ((threadIdx.x & 63) == 63) {
        s_mn[w] = (uint32_t)a;
        s_mx[w] = (uint32_t)b;
        s_bad[w] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m0 = s_mn[0], m1 = s_mx[0], bb = s_bad[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            m0 = min(m0, s_mn[k]);
            m1 = max(m1, s_mx[k]);
            bb |= s_bad[k];
        }
        part[0] = m0;
        part[1] = m1;
        part[2] = bb;
    }
}

// The raster's quantized min / max (ordered) and domain flag from the
// blocks' partials: one workgroup, its own loads of the partials all in
// flight per round, the three scalar words written once.
constexpr int MMF_T = 1024;
__global__ __launch_bounds__(MMF_T) void k_minmax_finish(const uint32_t* __restrict__ part, int nb,
                                                         uint32_t* __restrict__ scal) {
    uint32_t mn = 0xffffffffu, mx = 0u, bad = 0u;
#pragma unroll 8
    for (int i = threadIdx.x; i < nb; i += MMF_T) {
        mn = min(mn, part[3 * (size_t)i]);
        mx = max(mx, part[3 * (size_t)i + 1]);
        bad |= part[3 * (size_t)i + 2];
    }
    __shared__ uint32_t s_mn[MMF_T / 64], s_mx[MMF_T / 64], s_bad[MMF_T / 64];
    int a = (int)mn, b = (int)mx;
    CAGEO_SCAN_UP(dpp_i, a, b, (int)0xffffffffu, 0, minmax_after);
    const int w = threadIdx.x >> 6;
    const uint32_t wb = __any(bad != 0u) ? 1u : 0u;
    if ((threadIdx.x & 63) == 63) {
        s_mn[w] = (uint32_t)a;
        s_mx[w] = (uint32_t)b;
        s_bad[w] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m0 = s_mn[0], m1 = s_mx[0], bb = s_bad[0];
        for (int k = 1; k < MMF_T / 64; ++k) {
            m0 = min(m0, s_mn[k]);
            m1 = max(m1, s_mx[k]);
            bb |= s_bad[k];
        }
        scal[SC_MIN] = m0;
        scal[SC_MAX] = m1;
        scal[SC_BAD] = bb;
    }
}

__global__ __launch_bounds__(256) void k_quantize_minmax(float* __restrict__ z, int64_t n,
                                                         float nodata,
                                                         uint32_t* __restrict__ scal) {
    uint32_t mn = 0xffffffffu, mx = 0u, bad = 0u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float q = quantize_mm(z[i], nodata, &bad);
        z[i] = q;
        uint32_t o = float_to_ordered(q);
        mn = min(mn, o);
        mx = max(mx, o);
    }
    block_minmax(mn, mx, bad, scal);
}

// Fused realistic DEM + burn + quantisation + min/max for the Monte Carlo
// driver: one pass over dem, error field and stream flags.  Each block
// leaves its min/max partial in `part` (3 words per block) for
// k_minmax_finish, so no block waits on the global words.
__global__ __launch_bounds__(256) void k_realistic_burn_quant(
    const float* __restrict__ dem, const float* __restrict__ e, const uint8_t* __restrict__ sf,
    const int* __restrict__ off, const float* __restrict__ fac, int n_off, int R, int W, int H,
    float nodata_dem, float* __restrict__ zq, uint32_t* __restrict__ part) {
    float v[4];
    int gx[4], gy[4];
    burn_tile<true>(W, H, sf, off, fac, n_off, R, nodata_dem, dem, e, v, gx, gy);
    uint32_t mn = 0xffffffffu, mx = 0u, bad = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (gx[j] < W && gy[j] < H) {
            const float q = quantize_mm(v[j], nodata_dem, &bad);
            zq[(size_t)gy[j] * W + gx[j]] = q;
            const uint32_t o = float_to_ordered(q);
            mn = min(mn, o);
            mx = max(mx, o);
        }
    }
    block_minmax_store(mn, mx, bad, part + 3 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x));
}

inline unsigned grid_1d(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

}  // namespace

// Circular search array of StreamBurningAlgorithm.h:140-167 ordered by
// (d^2, y, x) as coordDistCompare (StreamBurningCommon.cpp:13-28), with the
// bevel factor sqrt(d^2)/R of StreamBurningAlgorithmCPUOpenMP.h:67-71
// evaluated in float on the host.
int prepare_offsets(Context* c, int R) {
    if (R < 0 || R > CAGEO_MAX_BEVEL_RADIUS) return CAGEO_ERR_ARG;
    if (c->off_radius == R) return CAGEO_OK;
    struct O {
        int d, x, y;
    };
    std::list<O> l;
    for (int y = -R; y <= R; ++y)
        for (int x = -R; x <= R; ++x)
            if ((x || y) && x * x + y * y <= R * R) l.push_back({x * x + y * y, x, y});
    l.sort([](const O& a, const O& b) {
        return a.d < b.d || (a.d == b.d && (a.y < b.y || (a.y == b.y && a.x < b.x)));
    });
    std::vector<int> off;
    std::vector<float> fac;
    for (const O& o : l) {
        off.push_back(o.x);
        off.push_back(o.y);
        float xd = (float)(-o.x), yd = (float)(-o.y);
        float dist = std::sqrt(xd * xd + yd * yd);
        fac.push_back(dist / (float)R);
    }
    c->n_offsets = (int)fac.size();
    if (c->n_offsets > 0) {
        CAGEO_HIP_TRY(hipMemcpyAsync(c->d_off, off.data(), off.size() * sizeof(int),
                                     hipMemcpyHostToDevice, c->stream));
        CAGEO_HIP_TRY(hipMemcpyAsync(c->d_factor, fac.data(), fac.size() * sizeof(float),
                                     hipMemcpyHostToDevice, c->stream));
        CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    c->off_radius = R;
    return CAGEO_OK;
}

int launch_stream_flags(Context* c, const float* stream, float nodata_stream, uint8_t* sflags,
                        int32_t* slabel) {
    hipLaunchKernelGGL(k_stream_flags, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, stream,
                       (int64_t)c->n, nodata_stream, sflags, slabel);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int launch_realistic(Context* c, const float* dem, float* field, float nodata) {
    hipLaunchKernelGGL(k_realistic, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, dem, field,
                       (int64_t)c->n, nodata);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int launch_burn(Context* c, const float* zin, float* zout, const uint8_t* sflags,
                float nodata_dem) {
    dim3 grid((c->W + BT_X - 1) / BT_X, (c->H + BT_Y - 1) / BT_Y);
    hipLaunchKernelGGL(k_burn, grid, dim3(256), 0, c->stream, zin, zout, sflags, c->d_off,
                       c->d_factor, c->n_offsets, c->off_radius, c->W, c->H, nodata_dem);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int launch_quantize(Context* c, float* dem, float nodata) {
    hipLaunchKernelGGL(k_quantize_minmax, dim3(grid_1d(c->n)), dim3(256), 0, c->stream, dem,
                       (int64_t)c->n, nodata, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

// Blocks of k_realistic_burn_quant (the size of the context's d_bmm, 3 words each).
size_t burn_blocks(int W, int H) {
    return (size_t)((W + BT_X - 1) / BT_X) * (size_t)((H + BT_Y - 1) / BT_Y);
}

// Writes the raster's min / max / domain words (SC_MIN, SC_MAX, SC_BAD) as
// well: they need no reset before it.
int launch_realistic_burn_quant(Context* c, const float* dem, const float* e, float* zq) {
    dim3 grid((c->W + BT_X - 1) / BT_X, (c->H + BT_Y - 1) / BT_Y);
    hipLaunchKernelGGL(k_realistic_burn_quant, grid, dim3(256), 0, c->stream, dem, e,
                       c->d_sflags, c->d_off, c->d_factor, c->n_offsets, c->off_radius, c->W, c->H,
                       c->nodata_dem, zq, c->d_bmm);
    hipLaunchKernelGGL(k_minmax_finish, dim3(1), dim3(MMF_T), 0, c->stream, c->d_bmm,
                       (int)burn_blocks(c->W, c->H), c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
