// Flow routing: FlowRoutingAlgorithm<T,U>::execute, FlowRoutingAlgorithm.h:
// 69-94; parity target FlowRoutingAlgorithm_CPU_OpenMP (D8 :53-151, flats
// :153-249 of FlowRoutingAlgorithmCPUOpenMP.h).
//
// D8: edge cells point out of the grid; cells at 0 are stream cells with no
// direction; otherwise the steepest of N, NE, E, SE, S, SW, W, NW with
// drop = z_c - z_n, diagonal drops divided by 1.41421356237309505f, strict
// '>' so the first of equal drops wins.  A cell with no positive drop is
// flat.
//
// Flats: the reference resolves them in rounds k = -1, 1, 2, ...: a flat c
// takes every neighbour n (same order, the last one wins) with
// flatDist[n] == k and z_n <= z_c.  That is a breadth-first distance
//   d(c) = 1 + min { d(n) : n an 8-neighbour, z_n <= z_c,
//                            n has a direction (d = 0) or is a flat },
// followed by direction = last n in order with d(n) = d(c) - 1 and
// z_n <= z_c.  d is computed here by monotone min-relaxation on LDS tiles.
#include "tiles.hpp"

namespace cageo {

namespace {

constexpr int D_INF = 0x3fffffff;
constexpr int D_STREAM = -2;
constexpr int D_OUT = -3;

__device__ constexpr int FDX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ constexpr int FDY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};

__global__ __launch_bounds__(256) void k_d8(const float* __restrict__ z,
                                            int8_t* __restrict__ fd,
                                            int32_t* __restrict__ dist, int W, int H) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    int dxi = 0, dyi = 0;
    if (x == 0) dxi = -1;
    else if (x == W - 1) dxi = 1;
    if (y == 0) dyi = -1;
    else if (y == H - 1) dyi = 1;
    int8_t ox = 0, oy = 0;
    int32_t d;
    if (dxi != 0 || dyi != 0) {
        ox = (int8_t)dxi;
        oy = (int8_t)dyi;
        d = 0;
    } else {
        const float zc = z[i];
        if (zc == 0.0f) {
            d = D_STREAM;
        } else {
            float steepest = 0.0f;
            int sk = -1;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                float s = zc - z[(size_t)(y + FDY[k]) * W + (x + FDX[k])];
                if (s > 0.0f) {
                    if (k & 1) s = s / 1.41421356237309505f;
                    if (s > steepest) {
                        steepest = s;
                        sk = k;
                    }
                }
            }
            if (sk >= 0) {
                ox = (int8_t)FDX[sk];
                oy = (int8_t)FDY[sk];
                d = 0;
            } else {
                d = D_INF;
            }
        }
    }
    reinterpret_cast<char2*>(fd)[i] = make_char2(ox, oy);
    dist[i] = d;
}

__global__ __launch_bounds__(256) void k_flat_bfs(const int* __restrict__ work,
                                                  const float* __restrict__ z,
                                                  int32_t* __restrict__ dist, int W, int H,
                                                  int ntx, int nty, int* __restrict__ work_out,
                                                  uint32_t* count_out,
                                                  uint32_t* __restrict__ mark, uint32_t epoch) {
    __shared__ int32_t dv[HALO_W * HALO_W];
    __shared__ float zv[HALO_W * HALO_W];
    const int t = work[blockIdx.x];
    const int tx = t % ntx, ty = t / ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    for (int k = threadIdx.x; k < HALO_W * HALO_W; k += TILE_THREADS) {
        int hy = k / HALO_W, hx = k - hy * HALO_W;
        int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        int32_t d = D_OUT;
        float zz = 0.0f;
        if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
            size_t gi = (size_t)gy * W + gx;
            d = dist[gi];
            zz = z[gi];
        }
        dv[k] = d;
        zv[k] = zz;
    }
    __syncthreads();

    const int lx = threadIdx.x & 63, strip = threadIdx.x >> 6;
    const int gx = x0 + lx;
    uint32_t flat = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
        if (dv[li] >= 1) flat |= 1u << k;
    }
    uint32_t changed_cells = 0;
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        if (flat) {
#pragma unroll
            for (int kk = 0; kk < STRIP; ++kk) {
                const int k = (sweep & 1) ? (STRIP - 1 - kk) : kk;
                if (!(flat & (1u << k))) continue;
                const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
                const float zc = zv[li];
                int best = dv[li];
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const int ni = li + FDY[d] * HALO_W + FDX[d];
                    const int dn = dv[ni];
                    if (dn >= 0 && zv[ni] <= zc && dn + 1 < best) best = dn + 1;
                }
                if (best < dv[li]) {
                    dv[li] = best;
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
            const int ly = strip * STRIP + k;
            dist[(size_t)(y0 + ly) * W + gx] = dv[(ly + 1) * HALO_W + lx + 1];
            bits |= border_bits(lx, ly);
        }
    }
    bits = block_or_bits(bits);
    if (threadIdx.x == 0 && bits)
        activate_tiles(bits, tx, ty, ntx, nty, work_out, count_out, mark, epoch);
}

__global__ __launch_bounds__(256) void k_flat_dir(const float* __restrict__ z,
                                                  const int32_t* __restrict__ dist,
                                                  int8_t* __restrict__ fd, int W, int H) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const int32_t d = dist[i];
    if (d < 1 || d >= D_INF) return;
    const float zc = z[i];
    int sk = -1;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int nx = x + FDX[k], ny = y + FDY[k];
        if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
        const size_t ni = (size_t)ny * W + nx;
        if (dist[ni] == d - 1 && z[ni] <= zc) sk = k;
    }
    if (sk >= 0) reinterpret_cast<char2*>(fd)[i] = make_char2((int8_t)FDX[sk], (int8_t)FDY[sk]);
}

}  // namespace

int flow_routing(Context* c, const float* z, int8_t* fd) {
    const int W = c->W, H = c->H;
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(k_d8, grid, dim3(256), 0, c->stream, z, fd, c->d_dist, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(worklist_all_tiles(c, 0));
    int passes = 0;
    CAGEO_TRY(run_tile_passes(
        c, 0, (uint32_t)c->ntiles, 1 << 20,
        [&](const int* win, uint32_t cnt, int* wout, uint32_t* slot, uint32_t epoch) {
            hipLaunchKernelGGL(k_flat_bfs, dim3(cnt), dim3(TILE_THREADS), 0, c->stream, win, z,
                               c->d_dist, W, H, c->ntx, c->nty, wout, slot, c->d_mark, epoch);
            CAGEO_HIP_TRY(hipGetLastError());
            return CAGEO_OK;
        },
        &passes));
    hipLaunchKernelGGL(k_flat_dir, grid, dim3(256), 0, c->stream, z, c->d_dist, fd, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
