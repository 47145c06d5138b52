// Fast mode of the 2-D (exponential) convolution: the symmetric filter in
// its low-rank separable form, f = sum_k lambda_k v_k v_k^T (filter.cpp,
// cageo_filter_low_rank), so
//   out[y][x] = sum_k lambda_k sum_j v_k[j] C_k[y][wrap(x - r + j)],
//   C_k[y][x] = sum_i v_k[i] in[wrap(y - r + i)][x].
// No reference counterpart: the reference runs the full 2-D sum
// (ConvolutionAlgorithmCPUOpenMP.h:94-102), which the exact mode keeps.
// This mode trades the bit-exact order for K x 2 x (2r+1) fused multiply-
// adds per cell instead of (2r+1)^2 products and sums; with the terms kept
// until the remainder is below 1e-6 of the filter it agrees with the exact
// convolution to well within 1e-5 relative.
//
// Column pass: 64 x 64 output tiles staged through LDS with their wrapped
// row halo; each thread owns a 16-row strip of one column and keeps the 16
// input rows its outputs need in a register ring (one LDS read per tap
// feeds 16 x KG fused multiply-adds).  Row pass: each thread owns 4 x 4
// outputs (4 columns as a float4); for every term the plane's tile with its
// wrapped column halo is staged in LDS and read as float4 pairs (8 values
// for 4 taps x 4 outputs), accumulating all terms in registers.
#include "common.hpp"

namespace cageo {

namespace {

__device__ inline int wrap_i(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

constexpr int LR_KG = 4;  // terms per column-pass launch slice

// taps: K rows of LC = (L + 15) & ~15 column taps (v_k, zero padded), so
// the tap loop runs whole ring turns; the LDS tile carries zero rows to
// match.
__global__ __launch_bounds__(256) void k_lr_cols(float* __restrict__ planes,
                                                 const float* __restrict__ in,
                                                 const float* __restrict__ taps, int K, int r,
                                                 int LC, int W, int H) {
    extern __shared__ float lds[];  // (64 + LC) rows x 64; rows from 64 + 2r on are zero
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
    const int k0 = blockIdx.z * LR_KG;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int rows = TILE + 2 * r, lrows = TILE + LC;
    const int x = x0 + tx;
    constexpr int B = 8;  // batches of independent loads
    for (int k0 = ty; k0 < lrows; k0 += 4 * B) {
        float v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = k0 + 4 * j;
            v[j] = 0.0f;
            if (k < rows && x < W) v[j] = in[(size_t)wrap_i(y0 - r + k, H) * W + x];
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (k0 + 4 * j < lrows) lds[(k0 + 4 * j) * TILE + tx] = v[j];
    }
    __syncthreads();
    float acc[LR_KG][STRIP];
#pragma unroll
    for (int g = 0; g < LR_KG; ++g)
#pragma unroll
        for (int j = 0; j < STRIP; ++j) acc[g][j] = 0.0f;
    const int base = ty * STRIP;
    // ring: at tap i, slot (i + j) % 16 holds input row base + i + j
    float w[STRIP];
#pragma unroll
    for (int j = 0; j < STRIP; ++j) w[j] = lds[(base + j) * TILE + tx];
    const int ng = min(LR_KG, K - k0);
    const float* tk[LR_KG];
#pragma unroll
    for (int g = 0; g < LR_KG; ++g) tk[g] = taps + (size_t)(k0 + (g < ng ? g : 0)) * LC;
    for (int ib = 0; ib < LC; ib += STRIP) {
#pragma unroll
        for (int u = 0; u < STRIP; ++u) {
            const int i = ib + u;
            float t[LR_KG];
#pragma unroll
            for (int g = 0; g < LR_KG; ++g) t[g] = g < ng ? tk[g][i] : 0.0f;
#pragma unroll
            for (int j = 0; j < STRIP; ++j) {
                const float v = w[(j + u) % STRIP];
#pragma unroll
                for (int g = 0; g < LR_KG; ++g) acc[g][j] = __builtin_fmaf(t[g], v, acc[g][j]);
            }
            // row base + i is done; its slot takes row base + i + 16
            w[u] = lds[(base + i + STRIP) * TILE + tx];
        }
    }
    if (x < W) {
#pragma unroll
        for (int g = 0; g < LR_KG; ++g) {
            if (g >= ng) continue;
            float* p = planes + (size_t)(k0 + g) * W * H;
#pragma unroll
            for (int j = 0; j < STRIP; ++j) {
                const int y = y0 + base + j;
                if (y < H) p[(size_t)y * W + x] = acc[g][j];
            }
        }
    }
}

// One 4-tap step of the row pass: loads the row segments of step q + 2 into
// `fut` and the taps of step q + 1, then adds the taps `t` of step q times
// the window (cur, nxt) into the 4 x 4 outputs (row a, column c).  On
// return `t` holds the taps of step q + 1.
__device__ __forceinline__ void lr_row_step(float (&acc)[4][4], const float4 (&cur)[4],
                                            const float4 (&nxt)[4], float4 (&fut)[4], float4& t,
                                            const float* row0, int RS, const float* T, int q) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
        fut[a] = *reinterpret_cast<const float4*>(row0 + a * RS + 4 * q + 8);
    const float4 t2 = *reinterpret_cast<const float4*>(T + 4 * q + 4);
    // keeps those loads issued ahead of the products: the scheduler would
    // otherwise sink them below to shorten their live ranges
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float win[8] = {cur[a].x, cur[a].y, cur[a].z, cur[a].w,
                              nxt[a].x, nxt[a].y, nxt[a].z, nxt[a].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float s = acc[a][c];
            s = __builtin_fmaf(t.x, win[c], s);
            s = __builtin_fmaf(t.y, win[c + 1], s);
            s = __builtin_fmaf(t.z, win[c + 2], s);
            s = __builtin_fmaf(t.w, win[c + 3], s);
            acc[a][c] = s;
        }
    }
    t = t2;
}

// rtaps: K rows of LP = (L + 3) & ~3 taps (lambda_k v_k, zero padded), then
// 4 zero floats.  TAPS_LDS: the taps are copied into LDS behind the tile, so
// that every operand load returns in order and the next step's loads stay
// in flight during the current step's products (scalar loads return out of
// order: a wave waits for all of them at each use).
template <bool TAPS_LDS>
__global__ __launch_bounds__(256) void k_lr_rows(float* __restrict__ out,
                                                 const float* __restrict__ planes,
                                                 const float* __restrict__ rtaps, int K, int r,
                                                 int LP, int S, int W, int H) {
    extern __shared__ float lds[];  // 64 rows x S (S >= 68 + LP, multiple of 4), taps
    // transposed grid (blockIdx.x walks tile rows): the left and right
    // neighbours, whose columns form this tile's halo, are gridDim.x blocks
    // away, on the same XCD and L2 when gridDim.x is a multiple of 8
    const int x0 = blockIdx.y * TILE, y0 = blockIdx.x * TILE;
    const int tx4 = threadIdx.x & 15, ty4 = threadIdx.x >> 4;
    const int cols = TILE + 2 * r;
    const float* taps = rtaps;
    if (TAPS_LDS) {
        float* tl = lds + TILE * S;
        for (int i = threadIdx.x; i < K * LP + 4; i += 256) tl[i] = rtaps[i];
        taps = tl;  // visible after the first barrier below
    }
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0f;
    const size_t n = (size_t)W * H;
    for (int k = 0; k < K; ++k) {
        const float* P = planes + (size_t)k * n;
        constexpr int B = 8;  // batches of independent loads
        for (int i0 = threadIdx.x; i0 < TILE * S; i0 += 256 * B) {
            float v[B];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int idx = i0 + 256 * j;
                v[j] = 0.0f;
                if (idx < TILE * S) {
                    const int row = idx / S, cc = idx - row * S;
                    const int gy = y0 + row;
                    if (cc < cols && gy < H) v[j] = P[(size_t)gy * W + wrap_i(x0 - r + cc, W)];
                }
            }
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (i0 + 256 * j < TILE * S) lds[i0 + 256 * j] = v[j];
        }
        __syncthreads();
        // rows ty4 + 16 a; each step takes one tap quadruple for all four
        const float* T = taps + (size_t)k * LP;
        const float* row0 = lds + ty4 * S + 4 * tx4;
        const int RS = 16 * S;
        float4 cur[4], nxt[4], fut[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            cur[a] = *reinterpret_cast<const float4*>(row0 + a * RS);
            nxt[a] = *reinterpret_cast<const float4*>(row0 + a * RS + 4);
        }
        float4 t = *reinterpret_cast<const float4*>(T);
        // three steps per trip: the roles of the three row-segment sets
        // rotate with period 3, so the rotation costs no register moves
        const int nq = LP >> 2;
        int q = 0;
        for (; q + 3 <= nq; q += 3) {
            lr_row_step(acc, cur, nxt, fut, t, row0, RS, T, q);
            lr_row_step(acc, nxt, fut, cur, t, row0, RS, T, q + 1);
            lr_row_step(acc, fut, cur, nxt, t, row0, RS, T, q + 2);
        }
        for (; q < nq; ++q) {
            lr_row_step(acc, cur, nxt, fut, t, row0, RS, T, q);
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                cur[a] = nxt[a];
                nxt[a] = fut[a];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int y = y0 + ty4 + 16 * a;
        if (y >= H) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int x = x0 + 4 * tx4 + c;
            if (x < W) out[(size_t)y * W + x] = acc[a][c];
        }
    }
}

}  // namespace

int lr_cols_geometry(int r, int* LC, size_t* lds) {
    *LC = (2 * r + 1 + 15) & ~15;
    *lds = (size_t)(TILE + *LC) * TILE * sizeof(float);
    return *lds <= LDS_LIMIT ? CAGEO_OK : CAGEO_ERR_ARG;
}

int lr_rows_geometry(int r, int* LP, int* S, size_t* lds) {
    const int L = 2 * r + 1;
    *LP = (L + 3) & ~3;
    *S = (TILE + *LP + 4 + 3) & ~3;  // float4 reads up to col 60 + LP + 7
    *lds = (size_t)TILE * (*S) * sizeof(float);
    return *lds <= LDS_LIMIT ? CAGEO_OK : CAGEO_ERR_ARG;
}

// Fast 2-D convolution of `field` into `out` (the low-rank form in c).
int launch_convolve_fast(Context* c, const float* field, float* out) {
    const int W = c->W, H = c->H, r = c->radius, K = c->lr_k;
    if (K < 1 || !c->d_lr_col || !c->d_lr_row || !c->d_lr_planes) return CAGEO_ERR_STATE;
    int LC, LP, S;
    size_t lds_c, lds_r;
    if (lr_cols_geometry(r, &LC, &lds_c) != CAGEO_OK ||
        lr_rows_geometry(r, &LP, &S, &lds_r) != CAGEO_OK)
        return CAGEO_ERR_ARG;
    dim3 gc((W + TILE - 1) / TILE, (H + TILE - 1) / TILE, (K + LR_KG - 1) / LR_KG);
    hipLaunchKernelGGL(k_lr_cols, gc, dim3(256), lds_c, c->stream, c->d_lr_planes, field,
                       c->d_lr_col, K, r, LC, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    dim3 gr((H + TILE - 1) / TILE, (W + TILE - 1) / TILE);  // transposed
    // the taps go into LDS behind the tile when both fit
    const size_t taps_bytes = ((size_t)K * LP + 4) * sizeof(float);
    if (lds_r + taps_bytes <= LDS_LIMIT)
        hipLaunchKernelGGL(k_lr_rows<true>, gr, dim3(256), lds_r + taps_bytes, c->stream, out,
                           c->d_lr_planes, c->d_lr_row, K, r, LP, S, W, H);
    else
        hipLaunchKernelGGL(k_lr_rows<false>, gr, dim3(256), lds_r, c->stream, out,
                           c->d_lr_planes, c->d_lr_row, K, r, LP, S, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
