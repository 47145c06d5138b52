// Monte Carlo check of a border-probability map, and refinement of the
// windows that fail it.
//
// No reference counterpart: the reference only runs the Monte Carlo
// (main.cpp:180-283).  A probability map q (from the surrogate network, or
// any other source) is judged against k realizations of the context's own
// Monte Carlo (cageo_run_realizations): with f = count / k per cell,
//   bias = mean(f - q)                          E = mean(p - q)
//   mse  = mean((f - q)^2 - f (1 - f) / (k - 1)) E = mean((p - q)^2)
// where p is the true border probability (E[f] = p, and f (1 - f) / (k - 1)
// is the unbiased estimate of the sampling variance p (1 - p) / k), so both
// are unbiased for any k >= 2 however few realizations are drawn.  A map is
// accepted when |bias| <= bias_tol and sqrt(max(mse, 0)) <= rms_tol.  The
// statistics are taken per window of `window` x `window` cells (ragged at the
// right and bottom edges; window <= 0: the whole raster as one window).
//
// Refinement: n realizations of the Monte Carlo over the whole DEM (the
// labels of a window depend on flow paths that leave it), whose frequency
// replaces the map in every window that failed.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace cageo {

namespace {

constexpr int CHK_THREADS = 256;
constexpr int CHK_MAX_BLOCKS = 1024;

struct WindowGrid {
    int win_w, win_h, nwx, nwy;
};

WindowGrid window_grid(const Context* c, int window) {
    WindowGrid g;
    g.win_w = window > 0 ? std::min(window, c->W) : c->W;
    g.win_h = window > 0 ? std::min(window, c->H) : c->H;
    g.nwx = (c->W + g.win_w - 1) / g.win_w;
    g.nwy = (c->H + g.win_h - 1) / g.win_h;
    return g;
}

// Block b sums rows s, s + S, s + 2S, ... of window b / S (s = b % S):
// part[b] = (sum f - q, sum (f - q)^2 - f (1 - f) / (k - 1), cells).
__global__ __launch_bounds__(CHK_THREADS) void k_check_windows(
    const float* __restrict__ q, const float* __restrict__ cnt, int W, int H, int win_w,
    int win_h, int nwx, int slices, int k, double* __restrict__ part) {
    __shared__ double red[3][CHK_THREADS];
    const int w = blockIdx.x / slices, s = blockIdx.x % slices;
    const int x0 = (w % nwx) * win_w, y0 = (w / nwx) * win_h;
    const int x1 = min(x0 + win_w, W), y1 = min(y0 + win_h, H);
    const double inv_k = 1.0 / k, inv_km1 = k > 1 ? 1.0 / (k - 1) : 0.0;
    double sb = 0.0, sm = 0.0, sn = 0.0;
    for (int y = y0 + s; y < y1; y += slices) {
        const size_t row = (size_t)y * W;
        for (int x = x0 + (int)threadIdx.x; x < x1; x += CHK_THREADS) {
            const double f = (double)cnt[row + x] * inv_k;
            const double d = f - (double)q[row + x];
            sb += d;
            sm += d * d - f * (1.0 - f) * inv_km1;
            sn += 1.0;
        }
    }
    red[0][threadIdx.x] = sb;
    red[1][threadIdx.x] = sm;
    red[2][threadIdx.x] = sn;
    __syncthreads();
    for (int off = CHK_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 3; ++j) part[(size_t)blockIdx.x * 3 + j] = red[j][0];
}

__global__ void k_zero(float* __restrict__ p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// prob = count / n in the cells of windows whose flag is 0.
__global__ void k_refine_windows(float* __restrict__ prob, const float* __restrict__ cnt, int W,
                                 int H, int win_w, int win_h, int nwx,
                                 const int* __restrict__ accepted, float inv_n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)W * H;
         i += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long long)y * W);
        if (!accepted[(y / win_h) * nwx + x / win_w]) prob[i] = cnt[i] * inv_n;
    }
}

}  // namespace

}  // namespace cageo

using cageo::Context;

extern "C" {

int cageo_check_windows(cageo_context* ctx, int window, int* nwx, int* nwy) {
    if (!ctx || !nwx || !nwy) return CAGEO_ERR_ARG;
    const cageo::WindowGrid g = cageo::window_grid(&ctx->c, window);
    *nwx = g.nwx;
    *nwy = g.nwy;
    return CAGEO_OK;
}

int cageo_check_probability_map(cageo_context* ctx, const float* d_prob, int k, uint64_t seed,
                                int window, double bias_tol, double rms_tol, float* d_scratch,
                                int* accepted, double* bias, double* rms) {
    if (!ctx || !d_prob || !d_scratch || k < 2 || k > 1000000 || !(bias_tol >= 0.0) ||
        !(rms_tol >= 0.0))
        return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    if (!c->mc_ready || c->model < 0) return CAGEO_ERR_STATE;
    const cageo::WindowGrid g = cageo::window_grid(c, window);
    const int nwin = g.nwx * g.nwy;
    const int slices = std::max(1, std::min(g.win_h, cageo::CHK_MAX_BLOCKS / std::max(1, nwin)));
    const size_t nblocks = (size_t)nwin * slices;
    // partial sums in the tracing workspace (8 bytes per cell), free between
    // realizations
    if (nblocks * 3 * sizeof(double) > c->n * sizeof(uint64_t)) return CAGEO_ERR_ARG;
    double* part = reinterpret_cast<double*>(c->d_ptr);
    hipLaunchKernelGGL(cageo::k_zero, dim3(1024), dim3(256), 0, c->stream, d_scratch,
                       (long long)c->n);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(cageo_run_realizations(ctx, d_scratch, k, seed, 0));
    hipLaunchKernelGGL(cageo::k_check_windows, dim3((unsigned)nblocks), dim3(cageo::CHK_THREADS),
                       0, c->stream, d_prob, d_scratch, c->W, c->H, g.win_w, g.win_h, g.nwx,
                       slices, k, part);
    CAGEO_HIP_TRY(hipGetLastError());
    std::vector<double> h(nblocks * 3);
    CAGEO_HIP_TRY(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    for (int w = 0; w < nwin; ++w) {
        double sb = 0.0, sm = 0.0, sn = 0.0;
        for (int s = 0; s < slices; ++s) {
            const double* p = &h[((size_t)w * slices + s) * 3];
            sb += p[0];
            sm += p[1];
            sn += p[2];
        }
        const double b = sn > 0 ? sb / sn : 0.0;
        const double r = sn > 0 ? std::sqrt(std::max(sm / sn, 0.0)) : 0.0;
        if (bias) bias[w] = b;
        if (rms) rms[w] = r;
        if (accepted) accepted[w] = std::fabs(b) <= bias_tol && r <= rms_tol;
    }
    return CAGEO_OK;
}

int cageo_refine_probability_map(cageo_context* ctx, float* d_prob, int window,
                                 const int* accepted, int n, uint64_t seed, float* d_scratch) {
    if (!ctx || !d_prob || !accepted || !d_scratch || n < 1) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    if (!c->mc_ready || c->model < 0) return CAGEO_ERR_STATE;
    const cageo::WindowGrid g = cageo::window_grid(c, window);
    const int nwin = g.nwx * g.nwy;
    bool all = true;
    for (int w = 0; w < nwin; ++w) all = all && accepted[w] != 0;
    if (all) return CAGEO_OK;
    hipLaunchKernelGGL(cageo::k_zero, dim3(1024), dim3(256), 0, c->stream, d_scratch,
                       (long long)c->n);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_TRY(cageo_run_realizations(ctx, d_scratch, n, seed, 0));
    // the window flags ride in the tracing workspace, free after the run
    int* flags = reinterpret_cast<int*>(c->d_ptr);
    if ((size_t)nwin * sizeof(int) > c->n * sizeof(*c->d_ptr)) return CAGEO_ERR_ARG;
    CAGEO_HIP_TRY(hipMemcpyAsync(flags, accepted, (size_t)nwin * sizeof(int),
                                 hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(cageo::k_refine_windows, dim3(2048), dim3(256), 0, c->stream, d_prob,
                       d_scratch, c->W, c->H, g.win_w, g.win_h, g.nwx, flags, 1.0f / (float)n);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    return CAGEO_OK;
}

}  // extern "C"
