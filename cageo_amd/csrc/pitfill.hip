// Pit filling: PitFillingAlgorithm<T>::execute, PitFillingAlgorithm.h:88-98,
// parity target performPitFilling, PitFillingAlgorithmCPUOpenMP.h:68-242.
//
// What the reference computes.  After quantising to mm (:84-106) it floods
// from the seeds (non-nodata edge cells and cells at 0) through buckets of
// 1/1000 m in increasing order (:162-237).  A popped cell c expands its
// unprocessed 8-neighbours n to max(z_n, spill_c); every cell is pushed once,
// into bucket toIndex(v) of its value v, and when popped it expands only if
// toElev(toIndex(v)) == v (:178-181).  Values that fail that round trip
// ("dead" values, e.g. 0.251 m) never expand, so with P the result
//   P(n) = max(Z(n), min { P(c) : c an 8-neighbour, reached, live(P(c)) }),
// seeds fixed at Z, unreached cells keep Z.
//
// How it is computed here, without a global priority queue.  Cells are
// FIXED (value known to equal P; live or dead) or OPEN.  Seeds start FIXED.
// Each round:
//   fill:    monotone relaxation from +inf of W(n) = max(Z(n), min relay(c))
//            over the OPEN cells, relay(c) = W(c) for OPEN and live FIXED
//            cells.  Every relay the reference makes is available, so at
//            convergence W <= P on the OPEN cells.  During the relaxation an
//            OPEN cell with W(n) == Z(n) and a live FIXED neighbour c with
//            P(c) <= Z(n) is fixed at once: P(n) <= max(Z(n), P(c)) = Z(n)
//            and P(n) >= Z(n).  A cell fixed dead stops relaying.
//   certify: after convergence an OPEN cell with a live FIXED neighbour c,
//            P(c) <= W(n), is fixed: P(n) <= max(Z(n), P(c)) <= W(n) <= P(n).
//   next:    OPEN cells restart at +inf.
// A round fixes at least the first OPEN cell the reference reaches (every
// relay path into it through OPEN cells has bottleneck >= its value), so a
// round that fixes nothing leaves only unreached cells; they keep Z.
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

// Loads relay (OPEN or live FIXED: W, else +inf) and crelay (live FIXED: W,
// else +inf) for the tile and its 1-cell halo.
__device__ inline void load_relays(float* relay, float* crelay, const float* __restrict__ w,
                                   const uint8_t* __restrict__ st, int x0, int y0, int W,
                                   int H) {
    for (int k = threadIdx.x; k < HALO_W * HALO_W; k += TILE_THREADS) {
        int hy = k / HALO_W, hx = k - hy * HALO_W;
        int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        float r = PF_INF, cr = PF_INF;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            uint8_t s = st[gi];
            if (s == PF_OPEN || s == PF_FIXED_LIVE) r = w[gi];
            if (s == PF_FIXED_LIVE) cr = r;
        }
        relay[k] = r;
        crelay[k] = cr;
    }
}

// Block-wide sum of per-thread counts, added to *dst by one thread.
__device__ inline void block_add(uint32_t v, uint32_t* dst) {
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
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

// Fill relaxation with immediate fixing of cells that sit at their own
// elevation next to a live fixed neighbour no higher than them.
__global__ __launch_bounds__(256) void k_pf_fill(
    const int* __restrict__ work, const float* __restrict__ zq, float* __restrict__ w,
    uint8_t* __restrict__ st, int W, int H, int ntx, int nty, int* __restrict__ work_out,
    uint32_t* count_out, uint32_t* __restrict__ mark, uint32_t epoch,
    uint32_t* __restrict__ topen, uint32_t* __restrict__ scal) {
    __shared__ float relay[HALO_W * HALO_W];
    __shared__ float crelay[HALO_W * HALO_W];
    const int t = work[blockIdx.x];
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    const float mn = min_elev_of(scal);
    load_relays(relay, crelay, w, st, x0, y0, W, H);

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

    uint32_t lowered = 0, fixed = 0, fixed_live = 0;
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        if (open) {
#pragma unroll
            for (int kk = 0; kk < STRIP; ++kk) {
                const int k = (sweep & 1) ? (STRIP - 1 - kk) : kk;
                if (!(open & (1u << k))) continue;
                const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
                const float m = min8(relay, li);
                const float nw = fmaxf(zr[k], fminf(wr[k], m));
                if (nw < wr[k]) {
                    wr[k] = nw;
                    relay[li] = nw;
                    changed = 1;
                    lowered |= 1u << k;
                }
                if (wr[k] == zr[k] && min8(crelay, li) <= wr[k]) {
                    open &= ~(1u << k);
                    fixed |= 1u << k;
                    changed = 1;
                    if (pf_live(wr[k], mn)) {
                        fixed_live |= 1u << k;
                        crelay[li] = wr[k];
                    } else {
                        relay[li] = PF_INF;
                    }
                }
            }
        }
        if (!__syncthreads_or(changed)) break;
    }

    uint32_t bits = 0, nfixed = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const size_t gi = (size_t)(y0 + ly) * W + gx;
        if (lowered & (1u << k)) w[gi] = wr[k];
        if (fixed & (1u << k)) {
            st[gi] = (fixed_live & (1u << k)) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
            ++nfixed;
        }
        if ((lowered | fixed) & (1u << k)) bits |= border_bits(lx, ly);
    }
    block_add(nfixed, &scal[SC_CERTIFIED]);
    const int any_open = __syncthreads_or(open != 0);
    bits = block_or_bits(bits);
    if (threadIdx.x == 0) {
        topen[t] = any_open ? 1u : 0u;
        if (bits) activate_tiles(bits, tx, ty, ntx, nty, work_out, count_out, mark, epoch);
    }
}

// Certification of the remaining OPEN cells after the fill has converged.
__global__ __launch_bounds__(256) void k_pf_cert(
    const int* __restrict__ work, const float* __restrict__ w, uint8_t* __restrict__ st, int W,
    int H, int ntx, int nty, int* __restrict__ work_out, uint32_t* count_out,
    uint32_t* __restrict__ mark, uint32_t epoch, uint32_t* __restrict__ topen,
    uint32_t* __restrict__ scal) {
    __shared__ float relay[HALO_W * HALO_W];
    __shared__ float crelay[HALO_W * HALO_W];
    const int t = work[blockIdx.x];
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    const float mn = min_elev_of(scal);
    load_relays(relay, crelay, w, st, x0, y0, W, H);

    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = x0 + lx;
    uint32_t open = 0;
    float wr[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int gy = y0 + strip * STRIP + k;
        wr[k] = PF_INF;
        if (gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            if (st[gi] == PF_OPEN) {
                wr[k] = w[gi];
                if (wr[k] < PF_INF) open |= 1u << k;
            }
        }
    }
    __syncthreads();

    uint32_t fixed = 0, fixed_live = 0;
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        if (open) {
#pragma unroll
            for (int kk = 0; kk < STRIP; ++kk) {
                const int k = (sweep & 1) ? (STRIP - 1 - kk) : kk;
                if (!(open & (1u << k))) continue;
                const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
                if (min8(crelay, li) <= wr[k]) {
                    open &= ~(1u << k);
                    fixed |= 1u << k;
                    changed = 1;
                    if (pf_live(wr[k], mn)) {
                        fixed_live |= 1u << k;
                        crelay[li] = wr[k];
                    }
                }
            }
        }
        if (!__syncthreads_or(changed)) break;
    }

    uint32_t bits = 0, nfixed = 0, still_open = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const int gy = y0 + ly;
        if (gx >= W || gy >= H) continue;
        const size_t gi = (size_t)gy * W + gx;
        if (fixed & (1u << k)) {
            st[gi] = (fixed_live & (1u << k)) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
            ++nfixed;
            if (fixed_live & (1u << k)) bits |= border_bits(lx, ly);
        } else if (st[gi] == PF_OPEN) {
            still_open = 1;
        }
    }
    block_add(nfixed, &scal[SC_CERTIFIED]);
    const int any_open = __syncthreads_or(still_open);
    bits = block_or_bits(bits);
    if (threadIdx.x == 0) {
        topen[t] = any_open ? 1u : 0u;
        if (bits) activate_tiles(bits, tx, ty, ntx, nty, work_out, count_out, mark, epoch);
    }
}

// Lists the tiles flagged as holding OPEN cells.
__global__ __launch_bounds__(256) void k_list_flagged(const uint32_t* __restrict__ topen,
                                                      int ntiles, int* __restrict__ work_out,
                                                      uint32_t* count_out) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x)
        if (topen[t]) work_out[atomicAdd(count_out, 1u)] = t;
}

// OPEN cells of the listed tiles restart from +inf.
__global__ __launch_bounds__(256) void k_pf_reset_open(const int* __restrict__ work,
                                                       float* __restrict__ w,
                                                       const uint8_t* __restrict__ st, int W,
                                                       int H, int ntx) {
    const int t = work[blockIdx.x];
    const int x0 = (t % ntx) * TILE, y0 = (t / ntx) * TILE;
    const int gx = x0 + (threadIdx.x & 63);
    if (gx >= W) return;
    for (int k = 0; k < STRIP; ++k) {
        const int gy = y0 + (threadIdx.x >> 6) * STRIP + k;
        if (gy >= H) break;
        const size_t gi = (size_t)gy * W + gx;
        if (st[gi] == PF_OPEN) w[gi] = PF_INF;
    }
}

// Cells never reached keep their quantized elevation.
__global__ __launch_bounds__(256) void k_pf_finalize(const int* __restrict__ work,
                                                     float* __restrict__ w,
                                                     const float* __restrict__ zq,
                                                     const uint8_t* __restrict__ st, int W,
                                                     int H, int ntx) {
    const int t = work[blockIdx.x];
    const int x0 = (t % ntx) * TILE, y0 = (t / ntx) * TILE;
    const int gx = x0 + (threadIdx.x & 63);
    if (gx >= W) return;
    for (int k = 0; k < STRIP; ++k) {
        const int gy = y0 + (threadIdx.x >> 6) * STRIP + k;
        if (gy >= H) break;
        const size_t gi = (size_t)gy * W + gx;
        if (st[gi] == PF_OPEN) w[gi] = zq[gi];
    }
}

}  // namespace

static int list_open_tiles(Context* c, int buf, uint32_t* tiles) {
    CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_COUNT0 + buf, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_list_flagged, dim3((c->ntiles + 255) / 256), dim3(256), 0, c->stream,
                       c->d_topen, c->ntiles, c->d_work[buf], c->d_scal + SC_COUNT0 + buf);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(read_scalars(c));
    *tiles = c->h_scal[SC_COUNT0 + buf];
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

    const int max_passes = 1 << 18;
    auto fill = [&](const int* win, uint32_t cnt, int* wout, uint32_t* slot, uint32_t epoch) {
        hipLaunchKernelGGL(k_pf_fill, dim3(cnt), dim3(TILE_THREADS), 0, c->stream, win, zq, w,
                           c->d_st, W, H, c->ntx, c->nty, wout, slot, c->d_mark, epoch,
                           c->d_topen, c->d_scal);
        CAGEO_HIP_TRY(hipGetLastError());
        return CAGEO_OK;
    };
    auto cert = [&](const int* win, uint32_t cnt, int* wout, uint32_t* slot, uint32_t epoch) {
        hipLaunchKernelGGL(k_pf_cert, dim3(cnt), dim3(TILE_THREADS), 0, c->stream, win, w,
                           c->d_st, W, H, c->ntx, c->nty, wout, slot, c->d_mark, epoch,
                           c->d_topen, c->d_scal);
        CAGEO_HIP_TRY(hipGetLastError());
        return CAGEO_OK;
    };

    uint32_t tiles = (uint32_t)c->ntiles;
    CAGEO_TRY(worklist_all_tiles(c, 0));
    for (int round = 0;; ++round) {
        if (round > (int)c->n + 1) return CAGEO_ERR_NOT_CONVERGED;
        CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_CERTIFIED, 0, sizeof(uint32_t), c->stream));
        if (round > 0) {
            hipLaunchKernelGGL(k_pf_reset_open, dim3(tiles), dim3(TILE_THREADS), 0, c->stream,
                               c->d_work[0], w, c->d_st, W, H, c->ntx);
            CAGEO_HIP_TRY(hipGetLastError());
        }
        CAGEO_TRY(run_tile_passes(c, 0, tiles, max_passes, fill, nullptr));
        CAGEO_TRY(list_open_tiles(c, 0, &tiles));
        if (tiles == 0) break;
        CAGEO_TRY(run_tile_passes(c, 0, tiles, max_passes, cert, nullptr));
        CAGEO_TRY(list_open_tiles(c, 0, &tiles));
        if (tiles == 0) break;
        if (c->h_scal[SC_CERTIFIED] == 0) {
            hipLaunchKernelGGL(k_pf_finalize, dim3(tiles), dim3(TILE_THREADS), 0, c->stream,
                               c->d_work[0], w, zq, c->d_st, W, H, c->ntx);
            CAGEO_HIP_TRY(hipGetLastError());
            break;
        }
    }
    return CAGEO_OK;
}

}  // namespace cageo
