// Convolution stage: ConvolutionAlgorithm<T>::execute, ConvolutionAlgorithm.h:
// 296-312, with the accumulation order of the reference's CPU loops
// (ConvolutionAlgorithmCPUOpenMP.h:105-248): sum starts at 0.0f and adds
// in[..] * f[k] for ascending taps (2-D: filter row outer, column inner),
// every product and sum rounded separately (this file is compiled with
// -ffp-contract=off), periodic wrap of the source index.
//
// Gaussian: columns pass into a buffer, then rows pass back into the field
// (ConvolutionAlgorithmCPUOpenMP.h:80-92).  Exponential: one 2-D pass,
// out of place (:94-102).
//
// Layout: output tiles of 64 x 64 cells staged through LDS with their
// wrapped halo; taps are wave-uniform and come through scalar loads.
#include "common.hpp"

namespace cageo {

namespace {

__device__ inline int wrap_index(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// Columns pass: out[y][x] = sum_fy in[wrap(y - r + fy)][x] * f[fy].
__global__ __launch_bounds__(256) void k_conv_cols(float* __restrict__ out,
                                                   const float* __restrict__ in,
                                                   const float* __restrict__ f, int r, int W,
                                                   int H) {
    extern __shared__ float lds[];  // (64 + 2r) rows x 64
    const int L = 2 * r + 1;
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * TILE;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int rows = TILE + 2 * r;
    const int x = x0 + tx;
    // batches of independent loads (each store waits only for its batch)
    constexpr int B = 8;
    for (int k0 = ty; k0 < rows; k0 += 4 * B) {
        float v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = k0 + 4 * j;
            v[j] = 0.0f;
            if (k < rows && x < W) v[j] = in[(size_t)wrap_index(y0 - r + k, H) * W + x];
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (k0 + 4 * j < rows) lds[(k0 + 4 * j) * TILE + tx] = v[j];
    }
    __syncthreads();
    float acc[STRIP];
#pragma unroll
    for (int j = 0; j < STRIP; ++j) acc[j] = 0.0f;
    const float* base = lds + (ty * STRIP) * TILE + tx;
    for (int fy = 0; fy < L; ++fy) {
        const float t = f[fy];
        const float* p = base + fy * TILE;
#pragma unroll
        for (int j = 0; j < STRIP; ++j) {
            float prod = p[j * TILE] * t;
            acc[j] = acc[j] + prod;
        }
    }
    if (x < W) {
#pragma unroll
        for (int j = 0; j < STRIP; ++j) {
            int y = y0 + ty * STRIP + j;
            if (y < H) out[(size_t)y * W + x] = acc[j];
        }
    }
}

// Rows pass: out[y][x] = sum_fx in[y][wrap(x - r + fx)] * f[fx].  The grid
// is transposed (blockIdx.x walks tile rows): blocks go round robin over the
// 8 XCDs, so the left and right neighbours of a tile, whose halos overlap it,
// are gridDim.x blocks away, on the same XCD and its L2 when gridDim.x is a
// multiple of 8.
__global__ __launch_bounds__(256) void k_conv_rows(float* __restrict__ out,
                                                   const float* __restrict__ in,
                                                   const float* __restrict__ f, int r, int W,
                                                   int H) {
    extern __shared__ float lds[];  // 64 rows x (64 + 2r)
    const int L = 2 * r + 1;
    const int S = TILE + 2 * r;
    const int x0 = blockIdx.y * TILE, y0 = blockIdx.x * TILE;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    constexpr int B = 8;  // batches of independent loads
    for (int i0 = threadIdx.x; i0 < TILE * S; i0 += 256 * B) {
        float v[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int idx = i0 + 256 * j;
            v[j] = 0.0f;
            if (idx < TILE * S) {
                const int row = idx / S, k = idx - row * S;
                const int gy = y0 + row;
                if (gy < H) v[j] = in[(size_t)gy * W + wrap_index(x0 - r + k, W)];
            }
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (i0 + 256 * j < TILE * S) lds[i0 + 256 * j] = v[j];
    }
    __syncthreads();
    float acc[STRIP];
#pragma unroll
    for (int j = 0; j < STRIP; ++j) acc[j] = 0.0f;
    const float* base = lds + (ty * STRIP) * S + tx;
    for (int fx = 0; fx < L; ++fx) {
        const float t = f[fx];
#pragma unroll
        for (int j = 0; j < STRIP; ++j) {
            float prod = base[j * S + fx] * t;
            acc[j] = acc[j] + prod;
        }
    }
    const int x = x0 + tx;
    if (x < W) {
#pragma unroll
        for (int j = 0; j < STRIP; ++j) {
            int y = y0 + ty * STRIP + j;
            if (y < H) out[(size_t)y * W + x] = acc[j];
        }
    }
}

// One step of the 2-D pass: loads the row segments of step b + 2 into `fut`
// and the taps of step b + 1, then adds the four taps `t` of step b times
// the window (cur, nxt) into the 4 x 4 outputs (row j, column k).  On
// return `t` holds the taps of step b + 1.
__device__ __forceinline__ void conv_2d_step(float (&acc)[4][4], const float4 (&cur)[4],
                                             const float4 (&nxt)[4], float4 (&fut)[4],
                                             float4& t, const float* row0, int S,
                                             const float* frow, int b) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
        fut[j] = *reinterpret_cast<const float4*>(row0 + j * S + 4 * b + 8);
    const float4 t2 = *reinterpret_cast<const float4*>(frow + 4 * b + 4);
    // keeps those loads issued ahead of the products: the scheduler would
    // otherwise sink them below to shorten their live ranges
    __builtin_amdgcn_sched_barrier(0);
    const float tap[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float win[8] = {cur[j].x, cur[j].y, cur[j].z, cur[j].w,
                              nxt[j].x, nxt[j].y, nxt[j].z, nxt[j].w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float prod = win[k + i] * tap[i];
                acc[j][k] = acc[j][k] + prod;
            }
    }
    t = t2;
}

// Full 2-D pass (exponential model).  Each thread owns a 4 x 4 block of
// outputs; for every filter row the 4 x 4 outputs slide over one LDS row
// segment read as float4, with the taps of that row held in scalar
// registers.  The filter is stored with a row stride `fs` (multiple of 4,
// zero padded) so a tap quadruple is one aligned load.
//
// TAPS_LDS: the filter is copied into LDS behind the tile.  Scalar loads
// return out of order, so a wave that takes its taps through them waits
// for every outstanding load before each use; LDS reads return in order,
// which lets the next step's row segment and taps be in flight while the
// current step computes (a CU holds one wave per SIMD at large radii, so
// nothing else hides that latency).  Without it (filters too large for
// LDS) the taps come through scalar loads.
template <int TY, bool TAPS_LDS>
__global__ __launch_bounds__(16 * TY) void k_conv_2d(float* __restrict__ out,
                                                     const float* __restrict__ in,
                                                     const float* __restrict__ f, int r, int fs,
                                                     int S, int W, int H) {
    extern __shared__ float lds[];  // (4*TY + 2r) rows x S, then L x fs taps
    constexpr int NT = 16 * TY;
    constexpr int OUT_H = 4 * TY;
    const int L = 2 * r + 1;
    const int x0 = blockIdx.x * TILE, y0 = blockIdx.y * OUT_H;
    const int rows = OUT_H + 2 * r;
    const int cols = TILE + 2 * r;
    for (int k = threadIdx.y * 16 + threadIdx.x; k < rows * S; k += NT) {
        int rr = k / S, cc = k - rr * S;
        float v = 0.0f;
        if (cc < cols) {
            int gy = wrap_index(y0 - r + rr, H);
            int gx = wrap_index(x0 - r + cc, W);
            v = in[(size_t)gy * W + gx];
        }
        lds[k] = v;
    }
    const float* taps = f;
    if (TAPS_LDS) {
        float* fl = lds + rows * S;
        for (int k = threadIdx.y * 16 + threadIdx.x; k < L * fs; k += NT) fl[k] = f[k];
        taps = fl;
    }
    __syncthreads();

    const int tx4 = threadIdx.x, ty4 = threadIdx.y;
    float acc[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[j][k] = 0.0f;

    const int nb = L >> 2;      // full tap quadruples
    const int rem = L & 3;      // 1 or 3 (L is odd)
    // Each step takes one tap quadruple for the four output rows (128
    // products and sums) while the next step's row segments and taps load.
    // Per accumulator the order is filter row, then ascending taps.
    for (int fy = 0; fy < L; ++fy) {
        const float* frow = taps + (size_t)fy * fs;
        const float* row0 = lds + (ty4 * 4 + fy) * S + tx4 * 4;
        float4 cur[4], nxt[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cur[j] = *reinterpret_cast<const float4*>(row0 + j * S);
            nxt[j] = *reinterpret_cast<const float4*>(row0 + j * S + 4);
        }
        float4 t = *reinterpret_cast<const float4*>(frow);
        float4 fut[4];
        // three steps per trip: the roles of the three row-segment sets
        // rotate with period 3, so the rotation costs no register moves
        int b = 0;
        for (; b + 3 <= nb; b += 3) {
            conv_2d_step(acc, cur, nxt, fut, t, row0, S, frow, b);
            conv_2d_step(acc, nxt, fut, cur, t, row0, S, frow, b + 1);
            conv_2d_step(acc, fut, cur, nxt, t, row0, S, frow, b + 2);
        }
        for (; b < nb; ++b) {
            conv_2d_step(acc, cur, nxt, fut, t, row0, S, frow, b);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                cur[j] = nxt[j];
                nxt[j] = fut[j];
            }
        }
        if (rem) {
            const float tap[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float win[8] = {cur[j].x, cur[j].y, cur[j].z, cur[j].w,
                                      nxt[j].x, nxt[j].y, nxt[j].z, nxt[j].w};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    if (i < rem) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            float prod = win[k + i] * tap[i];
                            acc[j][k] = acc[j][k] + prod;
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int y = y0 + ty4 * 4 + j;
        if (y >= H) continue;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int x = x0 + tx4 * 4 + k;
            if (x < W) out[(size_t)y * W + x] = acc[j][k];
        }
    }
}

}  // namespace

// Largest LDS allocation a workgroup may request on gfx950.

int conv_2d_config(int r, int* ty, int* S, size_t* lds_bytes) {
    // S: LDS row stride, covering 64 + 2r columns plus one float4 of slack
    // for the remainder read, rounded to a multiple of 4.
    const int s = ((TILE + 2 * r + 4) + 3) & ~3;
    const int tys[3] = {16, 8, 4};
    for (int t : tys) {
        size_t bytes = (size_t)(4 * t + 2 * r) * s * sizeof(float);
        if (bytes <= LDS_LIMIT) {
            *ty = t;
            *S = s;
            *lds_bytes = bytes;
            return CAGEO_OK;
        }
    }
    return CAGEO_ERR_ARG;
}

int launch_convolve_to(Context* c, float* field, float* buf, float** result) {
    const int W = c->W, H = c->H, r = c->radius;
    if (c->model < 0 || !c->d_filter) return CAGEO_ERR_STATE;
    dim3 grid((W + TILE - 1) / TILE, (H + TILE - 1) / TILE);
    if (c->model == CAGEO_FILTER_GAUSSIAN) {
        size_t lds = (size_t)(TILE + 2 * r) * TILE * sizeof(float);
        if (lds > LDS_LIMIT) return CAGEO_ERR_ARG;
        hipLaunchKernelGGL(k_conv_cols, grid, dim3(256), lds, c->stream, buf, field,
                           c->d_filter, r, W, H);
        CAGEO_HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_conv_rows, dim3(grid.y, grid.x), dim3(256), lds, c->stream, field, buf,
                           c->d_filter, r, W, H);
        CAGEO_HIP_TRY(hipGetLastError());
        *result = field;
        return CAGEO_OK;
    }
    if (c->conv_fast) {  // FFT or low-rank separable form (within 1e-5 of the exact order)
        if (c->fft_ready) CAGEO_TRY(launch_convolve_fft(c, field, buf));
        else CAGEO_TRY(launch_convolve_fast(c, field, buf));
        *result = buf;
        return CAGEO_OK;
    }
    int ty, S;
    size_t lds;
    CAGEO_TRY(conv_2d_config(r, &ty, &S, &lds));
    // a tuned row-block height, when it fits the LDS budget
    if (c->conv_ty == 16 || c->conv_ty == 8 || c->conv_ty == 4) {
        const size_t b = (size_t)(4 * c->conv_ty + 2 * r) * S * sizeof(float);
        if (b <= LDS_LIMIT) {
            ty = c->conv_ty;
            lds = b;
        }
    }
    const int fs = (2 * r + 1 + 3 + 3) & ~3;
    // the taps go into LDS behind the tile when both fit
    const size_t taps_bytes = (size_t)(2 * r + 1) * fs * sizeof(float);
    const bool taps_lds = lds + taps_bytes <= LDS_LIMIT;
    if (taps_lds) lds += taps_bytes;
    dim3 g2((W + TILE - 1) / TILE, (H + 4 * ty - 1) / (4 * ty));
#define CAGEO_CONV_2D(TY_, TL_)                                                              \
    hipLaunchKernelGGL((k_conv_2d<TY_, TL_>), g2, dim3(16, TY_), lds, c->stream, buf, field, \
                       c->d_filter, r, fs, S, W, H)
    switch (ty) {
        case 16:
            if (taps_lds) CAGEO_CONV_2D(16, true);
            else CAGEO_CONV_2D(16, false);
            break;
        case 8:
            if (taps_lds) CAGEO_CONV_2D(8, true);
            else CAGEO_CONV_2D(8, false);
            break;
        default:
            if (taps_lds) CAGEO_CONV_2D(4, true);
            else CAGEO_CONV_2D(4, false);
            break;
    }
#undef CAGEO_CONV_2D
    CAGEO_HIP_TRY(hipGetLastError());
    *result = buf;
    return CAGEO_OK;
}

int launch_convolve(Context* c, float* field, float* buf) {
    float* res = nullptr;
    CAGEO_TRY(launch_convolve_to(c, field, buf, &res));
    if (res != field)
        CAGEO_HIP_TRY(hipMemcpyAsync(field, res, c->n * sizeof(float), hipMemcpyDeviceToDevice,
                                     c->stream));
    return CAGEO_OK;
}

}  // namespace cageo
