// Fast mode of the 2-D (exponential) convolution on power-of-two grids: the
// periodic correlation computed through the discrete Fourier transform of
// the W x H torus.
//
// The reference's convolution wraps every source index once around the grid
// (ConvolutionAlgorithmCPUOpenMP.h:105-162), so for a radius within the grid
// it is the circular correlation
//   out[y][x] = sum_{dy,dx} f[dy][dx] in[(y + dy) mod H][(x + dx) mod W],
// i.e. OUT = IN . conj(F) in the frequency domain, with F the transform of
// the filter placed on the torus.  The exponential filter is even in each
// axis (to the last bits of its construction), so F is real and even and a
// quarter (W/2 + 1) x (H/2 + 1) of it is stored, with the 1 / (N H)
// normalisation of the two inverse transforms folded in.  No reference
// counterpart: the exact mode keeps the reference's summation order
// bit for bit; this mode agrees with it to a few 1e-7 of the field's
// largest magnitude (float32 transforms, twiddles rounded from double).
//
// Three passes over HBM, no transposes:
//   1. rows: each row's W reals as W/2 complex z[n] = x[2n] + i x[2n+1],
//      one complex FFT of length N = W/2 and the split into the row's
//      half spectrum X[0..N] (N + 1 complex per row);
//   2. columns: C adjacent spectrum columns per workgroup (C x H complex in
//      LDS), forward FFT over y, product with the filter's spectrum, inverse
//      FFT, in place.  Workgroups are numbered so that the ones an XCD runs
//      at once take consecutive column groups: their row segments fill whole
//      cache lines of that XCD's L2;
//   3. rows: the inverse split and an inverse complex FFT of length N.
// The FFTs run in LDS: forward transforms decimation in time from
// bit-reversed loads, inverse ones decimation in frequency to bit-reversed
// stores (so neither needs a reordering pass), four radix-2 stages per LDS
// round trip (16 values in registers; one twiddle load per stage, the other
// factors are 16th roots of unity), LDS indices padded by one slot per 32.
#include <cmath>
#include <thread>
#include <vector>

#include "common.hpp"

namespace cageo {

namespace {

constexpr int FFT_T = 512;      // threads per workgroup
constexpr int FFT_ELEMS = 16384;  // complex values per workgroup (C x N)

__device__ inline int lpad(int i) { return i + (i >> 5); }
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// exp(-2 pi i k / 16), k = 0..15
__device__ constexpr float ROOT16_RE[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
    0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508978f,
    0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f};
__device__ constexpr float ROOT16_IM[16] = {
    0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508978f,
    0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f,
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f};

// Twiddle exp(-+2 pi i q / P) for P = 2^lgP <= M = 2^lgM, from the table
// tw[k] = exp(-2 pi i k / M); conjugated for the inverse.
template <bool INV>
__device__ inline float2 twiddle(const float2* __restrict__ tw, int lgM, int q, int lgP) {
    const float2 w = tw[q << (lgM - lgP)];
    return INV ? conjf2(w) : w;
}
// W_{2^(t+1)}^k as a 16th root (t <= 3, k < 2^t), conjugated for the inverse.
template <bool INV>
__device__ inline float2 root(int t, int k) {
    const int m = k << (3 - t);
    return make_float2(ROOT16_RE[m], INV ? -ROOT16_IM[m] : ROOT16_IM[m]);
}

// The steps are inlined: as calls their frames set the kernels' registers
// (215 VGPRs against about 70 inlined).
// r consecutive radix-2 stages (R = 2^r values per unit in registers) over C
// sequences of length N = 2^L stored at sh + c * sp (padded).  DIT (forward,
// INV false): stages s .. s+r-1 with half sizes h .. h 2^(r-1), h = 2^s.
// DIF (inverse, INV true): the same stages in reverse order.  A unit holds
// positions base + m h, m < R, base = (u / h) R h + (u mod h).
template <int R, bool INV>
__device__ __forceinline__ void fft_step(float2* sh, int sp, int C, int L, int lgh, const float2* __restrict__ tw,
                         int lgM) {
    constexpr int r = R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : 1;
    const int h = 1 << lgh;
    const int per = 1 << (L - r);  // units per sequence
    for (int u = threadIdx.x; u < C * per; u += FFT_T) {
        const int c = u >> (L - r), uu = u & (per - 1);
        const int j = uu & (h - 1);
        const int base = ((uu >> lgh) << (lgh + r)) + j;
        float2* s = sh + (size_t)c * sp;
        float2 v[R];
#pragma unroll
        for (int m = 0; m < R; ++m) v[m] = s[lpad(base + m * h)];
        if (!INV) {
#pragma unroll
            for (int t = 0; t < r; ++t) {
                const float2 wj = twiddle<false>(tw, lgM, j, lgh + t + 1);  // W_P^j, P = h 2^(t+1)
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    if (m & (1 << t)) continue;
                    const int k = m & ((1 << t) - 1);
                    const float2 w = k == 0 ? wj : cmul(wj, root<false>(t, k));
                    const float2 a = v[m], b = cmul(v[m + (1 << t)], w);
                    v[m] = cadd(a, b);
                    v[m + (1 << t)] = csub(a, b);
                }
            }
        } else {
#pragma unroll
            for (int t = r - 1; t >= 0; --t) {
                const float2 wj = twiddle<true>(tw, lgM, j, lgh + t + 1);
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    if (m & (1 << t)) continue;
                    const int k = m & ((1 << t) - 1);
                    const float2 w = k == 0 ? wj : cmul(wj, root<true>(t, k));
                    const float2 a = v[m], b = v[m + (1 << t)];
                    v[m] = cadd(a, b);
                    v[m + (1 << t)] = cmul(csub(a, b), w);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < R; ++m) s[lpad(base + m * h)] = v[m];
    }
}

template <bool INV>
__device__ __forceinline__ void fft_dispatch(float2* sh, int sp, int C, int L, int lgh, int r,
                             const float2* __restrict__ tw, int lgM) {
    switch (r) {
        case 4: fft_step<16, INV>(sh, sp, C, L, lgh, tw, lgM); break;
        case 3: fft_step<8, INV>(sh, sp, C, L, lgh, tw, lgM); break;
        case 2: fft_step<4, INV>(sh, sp, C, L, lgh, tw, lgM); break;
        default: fft_step<2, INV>(sh, sp, C, L, lgh, tw, lgM); break;
    }
}

// Forward transform of C sequences loaded in bit-reversed order; natural
// order out.
__device__ __forceinline__ void fft_forward(float2* sh, int sp, int C, int L, const float2* __restrict__ tw,
                            int lgM) {
    for (int s = 0; s < L;) {
        const int r = min(4, L - s);
        fft_dispatch<false>(sh, sp, C, L, s, r, tw, lgM);
        __syncthreads();
        s += r;
    }
}
// Inverse (unnormalised) transform of C sequences in natural order;
// bit-reversed order out.
__device__ __forceinline__ void fft_inverse(float2* sh, int sp, int C, int L, const float2* __restrict__ tw,
                            int lgM) {
    for (int s = L; s > 0;) {
        const int r = min(4, s);
        s -= r;
        fft_dispatch<true>(sh, sp, C, L, s, r, tw, lgM);
        __syncthreads();
    }
}

__device__ inline int brev(int i, int L) { return (int)(__builtin_bitreverse32((uint32_t)i) >> (32 - L)); }

// Pass 1: RW rows per workgroup, real -> half spectrum (N + 1 per row).
__global__ __launch_bounds__(FFT_T) void k_fft_rows_fwd(const float* __restrict__ in,
                                                        float2* __restrict__ spec, int W, int H,
                                                        int RW, const float2* __restrict__ tw,
                                                        int lgM) {
    extern __shared__ float2 sh[];
    const int N = W >> 1, L = 31 - __builtin_clz(N), sp = N + (N >> 5);
    const int y0 = blockIdx.x * RW;
    for (int i = threadIdx.x; i < RW * N; i += FFT_T) {
        const int c = i >> L, n = i & (N - 1);
        float2 z = make_float2(0.0f, 0.0f);
        if (y0 + c < H) z = reinterpret_cast<const float2*>(in + (size_t)(y0 + c) * W)[n];
        sh[(size_t)c * sp + lpad(brev(n, L))] = z;
    }
    __syncthreads();
    fft_forward(sh, sp, RW, L, tw, lgM);
    const int lgW = L + 1;
    for (int i = threadIdx.x; i < RW * (N + 1); i += FFT_T) {
        const int c = i / (N + 1), k = i - c * (N + 1);
        if (y0 + c >= H) continue;
        const float2* s = sh + (size_t)c * sp;
        const float2 a = s[lpad(k & (N - 1))], b = conjf2(s[lpad((N - k) & (N - 1))]);
        const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
        // (a - b) / (2i)
        const float2 o = make_float2(0.5f * (a.y - b.y), -0.5f * (a.x - b.x));
        spec[(size_t)(y0 + c) * (N + 1) + k] = cadd(e, cmul(twiddle<false>(tw, lgM, k, lgW), o));
    }
}

// Pass 2: C spectrum columns per workgroup: forward FFT over y, product with
// the filter's quarter spectrum q[kx][min(ky, H - ky)], inverse FFT, in place.
__global__ __launch_bounds__(FFT_T) void k_fft_cols(float2* __restrict__ spec,
                                                    const float* __restrict__ q, int W, int H,
                                                    int C, int groups,
                                                    const float2* __restrict__ tw, int lgM) {
    extern __shared__ float2 sh[];
    const int N = W >> 1, L = 31 - __builtin_clz(H), sp = H + (H >> 5);
    // XCD-aware numbering: workgroup b runs on XCD b % 8; the workgroups an
    // XCD holds at once take consecutive column groups
    const int per_xcd = (groups + 7) / 8;
    const int g = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (g >= groups) return;
    const int kx0 = g * C;
    for (int i = threadIdx.x; i < C * H; i += FFT_T) {
        const int y = i / C, c = i - y * C;
        float2 v = make_float2(0.0f, 0.0f);
        if (kx0 + c <= N) v = spec[(size_t)y * (N + 1) + kx0 + c];
        sh[(size_t)c * sp + lpad(brev(y, L))] = v;
    }
    __syncthreads();
    fft_forward(sh, sp, C, L, tw, lgM);
    const int hq = (H >> 1) + 1;
    for (int i = threadIdx.x; i < C * H; i += FFT_T) {
        const int c = i >> L, ky = i & (H - 1);
        if (kx0 + c > N) continue;
        const float f = q[(size_t)(kx0 + c) * hq + min(ky, H - ky)];
        float2& v = sh[(size_t)c * sp + lpad(ky)];
        v = make_float2(v.x * f, v.y * f);
    }
    __syncthreads();
    fft_inverse(sh, sp, C, L, tw, lgM);
    for (int i = threadIdx.x; i < C * H; i += FFT_T) {
        const int y = i / C, c = i - y * C;
        if (kx0 + c <= N) spec[(size_t)y * (N + 1) + kx0 + c] = sh[(size_t)c * sp + lpad(brev(y, L))];
    }
}

// Pass 3: RW rows per workgroup, half spectrum -> real.
__global__ __launch_bounds__(FFT_T) void k_fft_rows_inv(const float2* __restrict__ spec,
                                                        float* __restrict__ out, int W, int H,
                                                        int RW, const float2* __restrict__ tw,
                                                        int lgM) {
    extern __shared__ float2 sh[];
    const int N = W >> 1, L = 31 - __builtin_clz(N), sp = N + (N >> 5), lgW = L + 1;
    const int y0 = blockIdx.x * RW;
    for (int i = threadIdx.x; i < RW * N; i += FFT_T) {
        const int c = i >> L, k = i & (N - 1);
        float2 z = make_float2(0.0f, 0.0f);
        if (y0 + c < H) {
            const float2* row = spec + (size_t)(y0 + c) * (N + 1);
            const float2 a = row[k], b = conjf2(row[N - k]);
            const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
            const float2 o = cmul(make_float2(0.5f * (a.x - b.x), 0.5f * (a.y - b.y)),
                                  twiddle<true>(tw, lgM, k, lgW));
            z = make_float2(e.x - o.y, e.y + o.x);  // e + i o
        }
        sh[(size_t)c * sp + lpad(k)] = z;
    }
    __syncthreads();
    fft_inverse(sh, sp, RW, L, tw, lgM);
    for (int i = threadIdx.x; i < RW * N; i += FFT_T) {
        const int c = i >> L, n = i & (N - 1);
        if (y0 + c < H)
            reinterpret_cast<float2*>(out + (size_t)(y0 + c) * W)[n] =
                sh[(size_t)c * sp + lpad(brev(n, L))];
    }
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int ilog2(int v) { return 31 - __builtin_clz((unsigned)v); }

size_t lds_bytes(int C, int N) { return (size_t)C * (N + (N >> 5)) * sizeof(float2); }

}  // namespace

bool fft_supported(int W, int H, int r) {
    return pow2(W) && pow2(H) && W >= 64 && H >= 32 && W <= 2 * FFT_ELEMS && H <= FFT_ELEMS &&
           r <= W / 2 && r <= H / 2;
}

void release_fft(Context* c) {
    if (c->d_fft_spec) (void)hipFree(c->d_fft_spec);
    if (c->d_fft_q) (void)hipFree(c->d_fft_q);
    if (c->d_fft_tw) (void)hipFree(c->d_fft_tw);
    c->d_fft_spec = nullptr;
    c->d_fft_q = nullptr;
    c->d_fft_tw = nullptr;
    c->fft_ready = false;
}

// The filter's spectrum (quarter, real, scaled) and the twiddle table, in
// double on the host.  Returns CAGEO_ERR_RANGE when the grid or the filter
// does not suit the FFT form (odd or too large sizes, a filter not even in
// both axes to 1e-6 of its largest tap): the low-rank form is used then.
int prepare_fft(Context* c) {
    release_fft(c);
    const int W = c->W, H = c->H, r = c->radius, L = 2 * r + 1;
    if (c->model != CAGEO_FILTER_EXPONENTIAL || c->h_taps.size() != (size_t)L * L ||
        !fft_supported(W, H, r))
        return CAGEO_ERR_RANGE;
    const std::vector<float>& f = c->h_taps;
    double fmax = 0.0, asym = 0.0;
    std::vector<double> fs((size_t)(r + 1) * (r + 1));  // even part, dy, dx >= 0
    for (int dy = 0; dy <= r; ++dy)
        for (int dx = 0; dx <= r; ++dx) {
            const double a = f[(size_t)(r + dy) * L + r + dx], b = f[(size_t)(r - dy) * L + r + dx];
            const double d = f[(size_t)(r + dy) * L + r - dx], e = f[(size_t)(r - dy) * L + r - dx];
            const double m = 0.25 * (a + b + d + e);
            fs[(size_t)dy * (r + 1) + dx] = m;
            fmax = std::max(fmax, std::max(std::max(std::fabs(a), std::fabs(b)),
                                           std::max(std::fabs(d), std::fabs(e))));
            asym = std::max(asym, std::max(std::max(std::fabs(a - m), std::fabs(b - m)),
                                           std::max(std::fabs(d - m), std::fabs(e - m))));
        }
    if (asym > 1e-6 * fmax) return CAGEO_ERR_RANGE;
    const int N = W / 2, hq = H / 2 + 1;
    const double two_pi = 6.283185307179586476925286766559;
    std::vector<double> cw(W), ch(H);
    for (int m = 0; m < W; ++m) cw[m] = std::cos(two_pi * m / W);
    for (int m = 0; m < H; ++m) ch[m] = std::cos(two_pi * m / H);
    // G[dy][kx] = sum_dx fs(dy, |dx|) cos(2 pi kx dx / W)
    std::vector<double> G((size_t)(r + 1) * (N + 1));
    for (int dy = 0; dy <= r; ++dy)
        for (int kx = 0; kx <= N; ++kx) {
            double s = fs[(size_t)dy * (r + 1)];
            for (int dx = 1; dx <= r; ++dx)
                s += 2.0 * fs[(size_t)dy * (r + 1) + dx] * cw[((size_t)kx * dx) % W];
            G[(size_t)dy * (N + 1) + kx] = s;
        }
    // q[kx][ky] = sum_dy G(|dy|, kx) cos(2 pi ky dy / H) / (N H)
    std::vector<float> q((size_t)(N + 1) * hq);
    const double scale = 1.0 / ((double)N * (double)H);
    const int nt = (int)std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int kx = t; kx <= N; kx += nt)
                for (int ky = 0; ky < hq; ++ky) {
                    double s = G[kx];
                    for (int dy = 1; dy <= r; ++dy)
                        s += 2.0 * G[(size_t)dy * (N + 1) + kx] * ch[((size_t)ky * dy) % H];
                    q[(size_t)kx * hq + ky] = (float)(s * scale);
                }
        });
    for (auto& x : th) x.join();
    const int M = std::max(W, H);
    std::vector<float2> tw(M);
    for (int k = 0; k < M; ++k)
        tw[k] = make_float2((float)std::cos(two_pi * k / M), (float)-std::sin(two_pi * k / M));
    if (hipMalloc(&c->d_fft_spec, (size_t)H * (N + 1) * sizeof(float2)) != hipSuccess ||
        hipMalloc(&c->d_fft_q, q.size() * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->d_fft_tw, tw.size() * sizeof(float2)) != hipSuccess) {
        release_fft(c);
        return CAGEO_ERR_NOMEM;
    }
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_fft_q, q.data(), q.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_fft_tw, tw.data(), tw.size() * sizeof(float2),
                                 hipMemcpyHostToDevice, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    c->fft_ready = true;
    return CAGEO_OK;
}

int launch_convolve_fft(Context* c, const float* field, float* out) {
    if (!c->fft_ready) return CAGEO_ERR_STATE;
    const int W = c->W, H = c->H, N = W / 2, lgM = ilog2(std::max(W, H));
    const int RW = std::max(1, FFT_ELEMS / std::max(N, 1));
    const int CC = std::max(1, FFT_ELEMS / H);
    const int groups = (N + 1 + CC - 1) / CC;
    const int blocks = ((groups + 7) / 8) * 8;
    hipLaunchKernelGGL(k_fft_rows_fwd, dim3((H + RW - 1) / RW), dim3(FFT_T), lds_bytes(RW, N),
                       c->stream, field, c->d_fft_spec, W, H, RW, c->d_fft_tw, lgM);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_fft_cols, dim3(blocks), dim3(FFT_T), lds_bytes(CC, H), c->stream,
                       c->d_fft_spec, c->d_fft_q, W, H, CC, groups, c->d_fft_tw, lgM);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_fft_rows_inv, dim3((H + RW - 1) / RW), dim3(FFT_T), lds_bytes(RW, N),
                       c->stream, c->d_fft_spec, out, W, H, RW, c->d_fft_tw, lgM);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
