// Pit filling: PitFillingAlgorithm<T>::execute, PitFillingAlgorithm.h:88-98,
// parity target performPitFilling, PitFillingAlgorithmCPUOpenMP.h:68-242.
//
// What the reference computes.  After quantising to mm (:84-106) it floods
// from the seeds (non-nodata edge cells and cells at 0) through buckets of
// 1/1000 m in increasing order (:162-237).  A popped cell c expands its
// unprocessed 8-neighbours n to max(z_n, spill_c); every cell is pushed once,
// into bucket toIndex(v) of its value v, and when popped it expands only if
// toElev(toIndex(v)) == v (:178-181).  Values that fail that round trip
// ("dead" values, e.g. 0.251 m) never expand, so
//   W(n) = max(Z(n), min { W(c) : c an 8-neighbour, reached, live(W(c)) }),
// seeds fixed at Z, unreached cells keep Z.
//
// How it is computed here, without a global priority queue:
//   round k: (1) fill: the greatest fixed point of
//                W(n) = max(Z(n), min relay(c)),  relay(c) = W(c) for OPEN
//                cells and live FIXED cells, +inf otherwise,
//            by monotone relaxation from +inf over the OPEN cells.  Every
//            relay of the reference is allowed, so W <= the reference's
//            value on every reached OPEN cell;
//            (2) certify: an OPEN cell n with a live FIXED neighbour c and
//                W(c) <= W(n) has the reference's value (the relay through c
//                bounds it from above, the fill from below); certification
//                propagates until closed;
//            (3) reset the still-OPEN cells to +inf and repeat.
// Each round certifies at least the first cell the reference reaches among
// the OPEN ones; a round that certifies nothing leaves only cells the
// reference never reaches, which keep their quantized elevation.
//
// Every comparison is on floats; the only arithmetic is the quantisation and
// the round-trip test, both evaluated exactly as the reference does.
#include "tiles.hpp"

namespace cageo {

namespace {

constexpr float PF_INF = __builtin_huge_valf();

__device__ inline float min8(const float* s, int i) {
    float a = fminf(s[i - HALO_W - 1], s[i - HALO_W]);
    float b = fminf(s[i - HALO_W + 1], s[i - 1]);
    float c = fminf(s[i + 1], s[i + HALO_W - 1]);
    float d = fminf(s[i + HALO_W], s[i + HALO_W + 1]);
    return fminf(fminf(a, b), fminf(c, d));
}

__device__ inline float min_elev_of(const uint32_t* scal) {
    return ordered_to_float(scal[SC_MIN]);
}

__global__ __launch_bounds__(256) void k_pf_init(const float* __restrict__ zq,
                                                 float* __restrict__ w,
                                                 uint8_t* __restrict__ st, int W, int H,
                                                 float nodata,
                                                 const uint32_t* __restrict__ scal) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const float z = zq[i];
    const float mn = min_elev_of(scal);
    uint8_t s;
    float v;
    if (z == nodata) {
        s = PF_NODATA;
        v = z;
    } else if (x == 0 || y == 0 || x == W - 1 || y == H - 1 || z == 0.0f) {
        // seeds, PitFillingAlgorithmCPUOpenMP.h:121-144
        s = pf_live(z, mn) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
        v = z;
    } else {
        s = PF_OPEN;
        v = PF_INF;
    }
    st[i] = s;
    w[i] = v;
}

// (1) fill relaxation over the OPEN cells of the listed tiles.
__global__ __launch_bounds__(256) void k_pf_fill(const int* __restrict__ work,
                                                 const float* __restrict__ zq,
                                                 float* __restrict__ w,
                                                 const uint8_t* __restrict__ st, int W, int H,
                                                 int ntx, int nty, int* __restrict__ work_out,
                                                 uint32_t* count_out,
                                                 uint32_t* __restrict__ mark, uint32_t epoch) {
    __shared__ float relay[HALO_W * HALO_W];
    const int t = work[blockIdx.x];
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;

    for (int k = threadIdx.x; k < HALO_W * HALO_W; k += TILE_THREADS) {
        int hy = k / HALO_W, hx = k - hy * HALO_W;
        int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        float v = PF_INF;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            uint8_t s = st[gi];
            if (s == PF_OPEN || s == PF_FIXED_LIVE) v = w[gi];
        }
        relay[k] = v;
    }

    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = x0 + lx;
    float zr[STRIP], wr[STRIP];
    uint32_t open = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        int gy = y0 + strip * STRIP + k;
        zr[k] = 0.0f;
        wr[k] = PF_INF;
        if (gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            if (st[gi] == PF_OPEN) {
                open |= 1u << k;
                zr[k] = zq[gi];
                wr[k] = w[gi];
            }
        }
    }
    __syncthreads();

    uint32_t changed_cells = 0;
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        if (open) {
#pragma unroll
            for (int kk = 0; kk < STRIP; ++kk) {
                const int k = (sweep & 1) ? (STRIP - 1 - kk) : kk;
                if (!(open & (1u << k))) continue;
                const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
                float m = min8(relay, li);
                float nw = fmaxf(zr[k], fminf(wr[k], m));
                if (nw < wr[k]) {
                    wr[k] = nw;
                    relay[li] = nw;
                    changed = 1;
                    changed_cells |= 1u << k;
                }
            }
        }
        if (!__syncthreads_or(changed)) break;
    }

    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if (changed_cells & (1u << k)) {
            int ly = strip * STRIP + k;
            w[(size_t)(y0 + ly) * W + gx] = wr[k];
            bits |= border_bits(lx, ly);
        }
    }
    bits = block_or_bits(bits);
    if (threadIdx.x == 0 && bits)
        activate_tiles(bits, tx, ty, ntx, nty, work_out, count_out, mark, epoch);
}

// (2) certification of OPEN cells through live FIXED neighbours.
__global__ __launch_bounds__(256) void k_pf_cert(const int* __restrict__ work,
                                                 const float* __restrict__ w,
                                                 uint8_t* __restrict__ st, int W, int H,
                                                 int ntx, int nty, int* __restrict__ work_out,
                                                 uint32_t* count_out,
                                                 uint32_t* __restrict__ mark, uint32_t epoch,
                                                 uint32_t* __restrict__ scal) {
    __shared__ float wv[HALO_W * HALO_W];
    __shared__ uint8_t sv[HALO_W * HALO_W];
    const int t = work[blockIdx.x];
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    const float mn = min_elev_of(scal);

    for (int k = threadIdx.x; k < HALO_W * HALO_W; k += TILE_THREADS) {
        int hy = k / HALO_W, hx = k - hy * HALO_W;
        int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        float v = PF_INF;
        uint8_t s = PF_NODATA;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            s = st[gi];
            v = w[gi];
        }
        wv[k] = v;
        sv[k] = s;
    }
    __syncthreads();

    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = x0 + lx;
    uint32_t open = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
        if (sv[li] == PF_OPEN && wv[li] < PF_INF) open |= 1u << k;
    }

    uint32_t newly = 0;
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        if (open) {
#pragma unroll
            for (int kk = 0; kk < STRIP; ++kk) {
                const int k = (sweep & 1) ? (STRIP - 1 - kk) : kk;
                if (!(open & (1u << k))) continue;
                const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
                const float v = wv[li];
                bool ok = false;
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const int ni = li + TDY[d] * HALO_W + TDX[d];
                    ok |= (sv[ni] == PF_FIXED_LIVE) && (wv[ni] <= v);
                }
                if (ok) {
                    sv[li] = pf_live(v, mn) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
                    open &= ~(1u << k);
                    newly |= 1u << k;
                    changed = 1;
                }
            }
        }
        if (!__syncthreads_or(changed)) break;
    }

    uint32_t bits = 0;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if (newly & (1u << k)) {
            const int ly = strip * STRIP + k;
            const int li = (ly + 1) * HALO_W + lx + 1;
            const uint8_t s = sv[li];
            st[(size_t)(y0 + ly) * W + gx] = s;
            ++cnt;
            if (s == PF_FIXED_LIVE) bits |= border_bits(lx, ly);
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, off);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&scal[SC_CERTIFIED], cnt);
    bits = block_or_bits(bits);
    if (threadIdx.x == 0 && bits)
        activate_tiles(bits, tx, ty, ntx, nty, work_out, count_out, mark, epoch);
}

// Lists the tiles holding OPEN cells (all tiles, one workgroup each), counts
// the OPEN cells and optionally resets them to +inf for the next round.
__global__ __launch_bounds__(256) void k_pf_open_tiles(float* __restrict__ w,
                                                       const uint8_t* __restrict__ st, int W,
                                                       int H, int ntx, int reset,
                                                       int* __restrict__ work_out,
                                                       uint32_t* count_out,
                                                       uint32_t* __restrict__ scal) {
    const int t = blockIdx.x;
    const int tx = t % ntx, ty = t / ntx;
    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = tx * TILE + lx;
    uint32_t cnt = 0;
    if (gx < W) {
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            const int gy = ty * TILE + strip * STRIP + k;
            if (gy >= H) break;
            const size_t gi = (size_t)gy * W + gx;
            if (st[gi] == PF_OPEN) {
                ++cnt;
                if (reset) w[gi] = PF_INF;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, off);
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&s_cnt, cnt);
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) {
        atomicAdd(&scal[SC_OPEN], s_cnt);
        uint32_t slot = atomicAdd(count_out, 1u);
        work_out[slot] = t;
    }
}

__global__ __launch_bounds__(256) void k_pf_finalize(float* __restrict__ w,
                                                     const float* __restrict__ zq,
                                                     const uint8_t* __restrict__ st, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (st[i] == PF_OPEN) w[i] = zq[i];
}

}  // namespace

// Lists tiles with OPEN cells into d_work[buf]; returns their count and the
// OPEN cell count through the scalar block.
static int list_open_tiles(Context* c, float* w, int reset, int buf, uint32_t* tiles,
                           uint32_t* open_cells) {
    CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_COUNT0 + buf, 0, sizeof(uint32_t), c->stream));
    CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_OPEN, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_pf_open_tiles, dim3(c->ntiles), dim3(TILE_THREADS), 0, c->stream, w,
                       c->d_st, c->W, c->H, c->ntx, reset, c->d_work[buf],
                       c->d_scal + SC_COUNT0 + buf, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(read_scalars(c));
    *tiles = c->h_scal[SC_COUNT0 + buf];
    *open_cells = c->h_scal[SC_OPEN];
    return CAGEO_OK;
}

// Pit fill of an already quantized raster whose min/max/bad scalars are in
// the scalar block (written by the quantisation kernel).
int pit_fill_quantized(Context* c, float* zq, float* w, float nodata, bool /*stats_ready*/) {
    CAGEO_TRY(read_scalars(c));
    if (c->h_scal[SC_BAD]) return CAGEO_ERR_RANGE;
    const float mn = ordered_to_float(c->h_scal[SC_MIN]);
    const float mx = ordered_to_float(c->h_scal[SC_MAX]);
    // The reference's bucket array holds fl(fl(max-min)*1000)+1 buckets; past
    // 2^24 the float arithmetic indexes beyond it (undefined there).
    const float span = (mx - mn) * 1000.0f;
    if (!(span < 16777216.0f)) return CAGEO_ERR_RANGE;

    const int W = c->W, H = c->H;
    hipLaunchKernelGGL(k_pf_init, dim3((W + 63) / 64, (H + 3) / 4), dim3(256), 0, c->stream,
                       zq, w, c->d_st, W, H, nodata, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());

    const int max_passes = 1 << 20;
    uint32_t tiles = 0, open_cells = 0;
    CAGEO_TRY(worklist_all_tiles(c, 0));
    tiles = (uint32_t)c->ntiles;
    for (int round = 0;; ++round) {
        if (round > 0) {
            CAGEO_TRY(list_open_tiles(c, w, 1, 0, &tiles, &open_cells));
        }
        int passes = 0;
        CAGEO_TRY(run_tile_passes(
            c, 0, tiles, max_passes,
            [&](const int* win, uint32_t cnt, int* wout, uint32_t* slot, uint32_t epoch) {
                hipLaunchKernelGGL(k_pf_fill, dim3(cnt), dim3(TILE_THREADS), 0, c->stream, win,
                                   zq, w, c->d_st, W, H, c->ntx, c->nty, wout, slot, c->d_mark,
                                   epoch);
                CAGEO_HIP_TRY(hipGetLastError());
                return CAGEO_OK;
            },
            &passes));

        CAGEO_TRY(list_open_tiles(c, w, 0, 0, &tiles, &open_cells));
        CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_CERTIFIED, 0, sizeof(uint32_t), c->stream));
        CAGEO_TRY(run_tile_passes(
            c, 0, tiles, max_passes,
            [&](const int* win, uint32_t cnt, int* wout, uint32_t* slot, uint32_t epoch) {
                hipLaunchKernelGGL(k_pf_cert, dim3(cnt), dim3(TILE_THREADS), 0, c->stream, win,
                                   w, c->d_st, W, H, c->ntx, c->nty, wout, slot, c->d_mark,
                                   epoch, c->d_scal);
                CAGEO_HIP_TRY(hipGetLastError());
                return CAGEO_OK;
            },
            &passes));
        CAGEO_TRY(read_scalars(c));
        const uint32_t certified = c->h_scal[SC_CERTIFIED];
        CAGEO_TRY(list_open_tiles(c, w, 0, 0, &tiles, &open_cells));
        if (open_cells == 0) break;
        if (certified == 0) {
            hipLaunchKernelGGL(k_pf_finalize, dim3((unsigned)std::min<size_t>((c->n + 255) / 256, 65536)),
                               dim3(256), 0, c->stream, w, zq, c->d_st, (int64_t)c->n);
            CAGEO_HIP_TRY(hipGetLastError());
            break;
        }
        if (round > max_passes) return CAGEO_ERR_NOT_CONVERGED;
    }
    return CAGEO_OK;
}

}  // namespace cageo
