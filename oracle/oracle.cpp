/*
 * CPU oracle for the Monte Carlo drainage-basin pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  The product never links, loads or calls this file.
 *
 * A clean-room restatement of the reference's OpenMP CPU stages
 * (jsoksane/cageo).  Each function cites the reference file:line it
 * follows; paths are relative to the reference's src/ directory.  The
 * reference itself cannot be compiled here (its CPU headers pull in the
 * CUDA runtime, Boost, Eigen, FFTW and sigar), so there is no compiled
 * reference to compare against; see the pinning notes in each function.
 *
 * Compile with -ffp-contract=off: the reference's Makefile builds the CPU
 * code without FMA contraction, and bit parity of the convolution depends
 * on every product and sum being rounded separately.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <queue>
#include <vector>

#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Random field: Philox4x32-10 + Box-Muller in double.                       */
/* ------------------------------------------------------------------------ */

static inline uint32_t mulhi32(uint32_t a, uint32_t b, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *lo = (uint32_t)p;
    return (uint32_t)(p >> 32);
}

/* Philox4x32-10: 10 rounds of the 4x32 S-box with multipliers
 * 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
 * 0xBB67AE85 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Pinned by the published known-answer
 * vectors in tests/test_oracle_rng.py. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0, lo1;
        uint32_t hi0 = mulhi32(0xD2511F53u, c0, &lo0);
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c2, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ln(u) for u in (0, 1]: u = m 2^e with m in (sqrt(1/2), sqrt(2)],
 * ln m = 2 atanh(s), s = (m-1)/(m+1), atanh by its Taylor series
 * s + s^3/3 + ... + s^21/21 (|s| <= 0.1716, truncation < 1e-17).
 * Every operation is a correctly rounded IEEE double op, so the device
 * copy of this routine gives identical bits. */
double or_rng_log(double u) {
    uint64_t bits;
    std::memcpy(&bits, &u, 8);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    uint64_t mb = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    std::memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double q = 1.0 / 21.0;
    q = q * z + 1.0 / 19.0;
    q = q * z + 1.0 / 17.0;
    q = q * z + 1.0 / 15.0;
    q = q * z + 1.0 / 13.0;
    q = q * z + 1.0 / 11.0;
    q = q * z + 1.0 / 9.0;
    q = q * z + 1.0 / 7.0;
    q = q * z + 1.0 / 5.0;
    q = q * z + 1.0 / 3.0;
    double two_s = 2.0 * s;
    double lnm = two_s + two_s * (z * q);
    return (double)e * 0.6931471805599453 + lnm;
}

/* cos(2 pi t), sin(2 pi t) for t in [0, 1): x = 4t quarter turns, reduced
 * exactly to r = x - round(x) in [-1/2, 1/2], a = r pi/2 in [-pi/4, pi/4],
 * then the Taylor series of sin (to a^17) and cos (to a^18). */
void or_rng_cossin_2pi(double t, double* c, double* s) {
    double x = 4.0 * t;
    double q = std::floor(x + 0.5);
    double r = x - q;
    double a = r * 1.5707963267948966;
    double a2 = a * a;
    double ps = -1.0 / 355687428096000.0;          /* -1/17! */
    ps = ps * a2 + 1.0 / 1307674368000.0;          /* 1/15! */
    ps = ps * a2 - 1.0 / 6227020800.0;             /* -1/13! */
    ps = ps * a2 + 1.0 / 39916800.0;               /* 1/11! */
    ps = ps * a2 - 1.0 / 362880.0;                 /* -1/9! */
    ps = ps * a2 + 1.0 / 5040.0;                   /* 1/7! */
    ps = ps * a2 - 1.0 / 120.0;                    /* -1/5! */
    ps = ps * a2 + 1.0 / 6.0;                      /* 1/3! */
    double sin_a = a - a * (a2 * ps);
    double pc = 1.0 / 6402373705728000.0;          /* 1/18! */
    pc = pc * a2 - 1.0 / 20922789888000.0;         /* -1/16! */
    pc = pc * a2 + 1.0 / 87178291200.0;            /* 1/14! */
    pc = pc * a2 - 1.0 / 479001600.0;              /* -1/12! */
    pc = pc * a2 + 1.0 / 3628800.0;                /* 1/10! */
    pc = pc * a2 - 1.0 / 40320.0;                  /* -1/8! */
    pc = pc * a2 + 1.0 / 720.0;                    /* 1/6! */
    pc = pc * a2 - 1.0 / 24.0;                     /* -1/4! */
    pc = pc * a2 + 0.5;
    double cos_a = 1.0 - a2 * pc;
    int qi = ((int)q) & 3;
    switch (qi) {
        case 0: *c = cos_a; *s = sin_a; break;
        case 1: *c = -sin_a; *s = cos_a; break;
        case 2: *c = -cos_a; *s = -sin_a; break;
        default: *c = sin_a; *s = -cos_a; break;
    }
}

static inline void box_muller(uint32_t a, uint32_t b, double* n0, double* n1) {
    double u1 = (double)((a >> 8) + 1u) * (1.0 / 16777216.0);
    double u2 = (double)(b >> 8) * (1.0 / 16777216.0);
    double rad = std::sqrt(-2.0 * or_rng_log(u1));
    double c, s;
    or_rng_cossin_2pi(u2, &c, &s);
    *n0 = rad * c;
    *n1 = rad * s;
}

void or_random_field(float* field, int64_t n, uint64_t seed, uint32_t realization,
                     uint32_t stream_id, float mean, float std_dev) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int64_t nblocks = (n + 3) / 4;
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nblocks; ++b) {
        uint32_t ctr[4] = {(uint32_t)b, realization, stream_id, (uint32_t)(b >> 32)};
        uint32_t x[4];
        or_philox4x32_10(ctr, key, x);
        double g[4];
        box_muller(x[0], x[1], &g[0], &g[1]);
        box_muller(x[2], x[3], &g[2], &g[3]);
        for (int k = 0; k < 4; ++k) {
            int64_t i = 4 * b + k;
            if (i < n) {
                float r = (float)g[k];
                float scaled = r * std_dev;
                field[i] = mean + scaled;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Filter construction: ConvolutionAlgorithm.cpp:18-172.                     */
/* ------------------------------------------------------------------------ */

/* createCorrelogram, ConvolutionAlgorithm.cpp:49-82.  The exponential model
 * evaluates sqrt/exp in float (the unqualified calls resolve to the float
 * overloads), the Gaussian one in double (pow(float, int) promotes). */
static void correlogram(int model, float pr, float cell, int radius, float* out) {
    int i = 0;
    for (int y = -radius; y <= radius; ++y)
        for (int x = -radius; x <= radius; ++x, ++i) {
            if (model == 0) {
                float d = std::sqrt((float)(x * x + y * y));
                float scale = (pr / 3.0f) / cell;
                out[i] = (float)std::exp(-d / scale);
            } else {
                double den = std::pow((double)((pr / 1.73205080757f) / cell), 2);
                out[i] = (float)std::exp(-(double)(x * x + y * y) / den);
            }
        }
}

/* createFilterFromCorrelogram, ConvolutionAlgorithm.cpp:84-148:
 * K = shift(IDFT(|DFT(C)|^(1/2))) with the 2-D DFT evaluated directly in
 * double (FFTW 3.3.3 in the reference: float taps agree to rounding, exact
 * bits "parity unpinned").  Row-column DFT with exact twiddle indexing. */
static void filter_from_correlogram(const float* corr, int L, float* filter) {
    std::vector<double> cs(L), sn(L);
    for (int k = 0; k < L; ++k) {
        double ang = 2.0 * M_PI * (double)k / (double)L;
        cs[k] = std::cos(ang);
        sn[k] = std::sin(ang);
    }
    const size_t LL = (size_t)L * L;
    std::vector<double> rre(LL), rim(LL), fre(LL), fim(LL);
    /* rows: R[y][kx] = sum_x c[y][x] e^{-i 2pi kx x / L} */
    for (int y = 0; y < L; ++y)
        for (int kx = 0; kx < L; ++kx) {
            double re = 0.0, im = 0.0;
            for (int x = 0; x < L; ++x) {
                int t = (int)(((long)kx * x) % L);
                double v = (double)corr[y * L + x];
                re += v * cs[t];
                im -= v * sn[t];
            }
            rre[y * L + kx] = re;
            rim[y * L + kx] = im;
        }
    /* columns: F[ky][kx] = sum_y R[y][kx] e^{-i 2pi ky y / L} */
    for (int ky = 0; ky < L; ++ky)
        for (int kx = 0; kx < L; ++kx) {
            double re = 0.0, im = 0.0;
            for (int y = 0; y < L; ++y) {
                int t = (int)(((long)ky * y) % L);
                double a = rre[y * L + kx], b = rim[y * L + kx];
                re += a * cs[t] + b * sn[t];
                im += b * cs[t] - a * sn[t];
            }
            fre[ky * L + kx] = std::sqrt(std::sqrt(re * re + im * im));
        }
    /* inverse: K[y][x] = (1/L^2) Re sum_ky sum_kx M e^{+i 2pi (ky y + kx x)/L} */
    for (int ky = 0; ky < L; ++ky)
        for (int x = 0; x < L; ++x) {
            double re = 0.0, im = 0.0;
            for (int kx = 0; kx < L; ++kx) {
                int t = (int)(((long)kx * x) % L);
                double m = fre[ky * L + kx];
                re += m * cs[t];
                im += m * sn[t];
            }
            rre[ky * L + x] = re;
            rim[ky * L + x] = im;
        }
    for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x) {
            double re = 0.0;
            for (int ky = 0; ky < L; ++ky) {
                int t = (int)(((long)ky * y) % L);
                re += rre[ky * L + x] * cs[t] - rim[ky * L + x] * sn[t];
            }
            fim[y * L + x] = re / (double)LL;
        }
    const int shift = L / 2;
    for (int j = 0; j < L; ++j) {
        int nj = (j + shift) % L;
        for (int i = 0; i < L; ++i) {
            int ni = (i + shift) % L;
            filter[nj * L + ni] = (float)fim[j * L + i];
        }
    }
}

/* createKernelVector, ConvolutionAlgorithm.cpp:150-172: u0 * sqrt(s0) of
 * the float 2-D filter.  Eigen's JacobiSVD is replaced by power iteration
 * on K^T K (the filter is numerically rank one); the sign of u0 is not
 * fixed by the reference ("parity unpinned"), this build takes the
 * centre tap positive. */
static void kernel_vector(const float* k2, int L, float* out) {
    std::vector<double> A((size_t)L * L), v(L, 1.0 / std::sqrt((double)L)), w(L), u(L);
    for (size_t i = 0; i < (size_t)L * L; ++i) A[i] = (double)k2[i];
    for (int it = 0; it < 100; ++it) {
        for (int j = 0; j < L; ++j) {
            double s = 0.0;
            for (int i = 0; i < L; ++i) s += A[j * L + i] * v[i];
            u[j] = s;
        }
        for (int i = 0; i < L; ++i) {
            double s = 0.0;
            for (int j = 0; j < L; ++j) s += A[j * L + i] * u[j];
            w[i] = s;
        }
        double nrm = 0.0;
        for (int i = 0; i < L; ++i) nrm += w[i] * w[i];
        nrm = std::sqrt(nrm);
        for (int i = 0; i < L; ++i) v[i] = w[i] / nrm;
    }
    double s0 = 0.0;
    for (int j = 0; j < L; ++j) {
        double s = 0.0;
        for (int i = 0; i < L; ++i) s += A[j * L + i] * v[i];
        u[j] = s;
        s0 += s * s;
    }
    s0 = std::sqrt(s0);
    double sign = (u[L / 2] < 0.0) ? -1.0 : 1.0;
    for (int j = 0; j < L; ++j) out[j] = (float)(sign * (u[j] / s0) * std::sqrt(s0));
}

int or_create_filter(int model, float practical_range, float cell_size, int radius,
                     float* out) {
    if (radius < 0 || (model != 0 && model != 1) || !out) return OR_ERR_ARG;
    const int L = 2 * radius + 1;
    std::vector<float> corr((size_t)L * L), f2((size_t)L * L);
    correlogram(model, practical_range, cell_size, radius, corr.data());
    filter_from_correlogram(corr.data(), L, f2.data());
    if (model == 1)
        kernel_vector(f2.data(), L, out);
    else
        std::memcpy(out, f2.data(), sizeof(float) * L * L);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Convolution: ConvolutionAlgorithmCPUOpenMP.h:70-248.                      */
/* ------------------------------------------------------------------------ */

/* performConvolutionCols_OpenMP, ConvolutionAlgorithmCPUOpenMP.h:207-248 */
static void conv_cols(float* out, const float* in, const float* f, int r, int W, int H) {
    const int L = 2 * r + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum = 0.0f;
            for (int fy = 0; fy < L; ++fy) {
                int sy = y - r + fy;
                if (sy < 0) sy += H;
                else if (sy >= H) sy -= H;
                sum += in[(size_t)sy * W + x] * f[fy];
            }
            out[(size_t)y * W + x] = sum;
        }
}

/* performConvolutionRows_OpenMP, ConvolutionAlgorithmCPUOpenMP.h:164-205 */
static void conv_rows(float* out, const float* in, const float* f, int r, int W, int H) {
    const int L = 2 * r + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum = 0.0f;
            for (int fx = 0; fx < L; ++fx) {
                int sx = x - r + fx;
                if (sx < 0) sx += W;
                else if (sx >= W) sx -= W;
                sum += in[(size_t)y * W + sx] * f[fx];
            }
            out[(size_t)y * W + x] = sum;
        }
}

/* performConvolution_OpenMP, ConvolutionAlgorithmCPUOpenMP.h:105-162 */
static void conv_2d(float* out, const float* in, const float* f, int r, int W, int H) {
    const int L = 2 * r + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float sum = 0.0f;
            for (int fy = 0; fy < L; ++fy) {
                int sy = y - r + fy;
                if (sy < 0) sy += H;
                else if (sy >= H) sy -= H;
                const float* row = in + (size_t)sy * W;
                const float* frow = f + (size_t)fy * L;
                for (int fx = 0; fx < L; ++fx) {
                    int sx = x - r + fx;
                    if (sx < 0) sx += W;
                    else if (sx >= W) sx -= W;
                    sum += row[sx] * frow[fx];
                }
            }
            out[(size_t)y * W + x] = sum;
        }
}

/* performConvolution, ConvolutionAlgorithmCPUOpenMP.h:70-103.  The
 * reference wraps an index once (sx < 0 -> +W), which equals the modulo
 * only when radius <= width and radius <= height; larger radii read out
 * of bounds there and are rejected here. */
int or_convolve(float* field, int W, int H, const float* filter, int r, int model) {
    if (W <= 0 || H <= 0 || r < 0 || r > W || r > H) return OR_ERR_ARG;
    std::vector<float> buf(field, field + (size_t)W * H);
    if (model == 1) {
        conv_cols(buf.data(), field, filter, r, W, H);
        conv_rows(field, buf.data(), filter, r, W, H);
    } else if (model == 0) {
        conv_2d(buf.data(), field, filter, r, W, H);
        std::memcpy(field, buf.data(), sizeof(float) * (size_t)W * H);
    } else {
        return OR_ERR_ARG;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Realistic DEM: RealisticDemAlgorithm.h:49-70,                             */
/* AddSurfacesAlgorithmCPUOpenMP.h:36-56,                                    */
/* RemoveNegativeElevationsAlgorithmCPUOpenMP.h:34-48.                       */
/* ------------------------------------------------------------------------ */

void or_realistic_dem(const float* dem, float* field, int64_t n, float nodata) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float d = dem[i];
        if (d != nodata) field[i] = d + field[i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float v = field[i];
        if (v <= 0.0 && v != nodata) field[i] = 1.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* Stream burning.                                                           */
/* ------------------------------------------------------------------------ */

struct Off {
    int d2;
    int x, y;
};

/* generateCircularSearchArray, StreamBurningAlgorithm.h:140-167, with the
 * order of coordDistCompare, StreamBurningCommon.cpp:13-28 (d^2, then y,
 * then x).  The reference terminates the walk at an uninitialised (0,0)
 * slot; the offset count is explicit here. */
static std::vector<Off> circular_offsets(int R) {
    std::list<Off> l;
    for (int y = -R; y <= R; ++y)
        for (int x = -R; x <= R; ++x) {
            if (x == 0 && y == 0) continue;
            int d2 = x * x + y * y;
            if (d2 <= R * R) l.push_back({d2, x, y});
        }
    l.sort([](const Off& a, const Off& b) {
        if (a.d2 != b.d2) return a.d2 < b.d2;
        if (a.y != b.y) return a.y < b.y;
        return a.x < b.x;
    });
    return std::vector<Off>(l.begin(), l.end());
}

/* performStreamBurning, StreamBurningAlgorithmCPUOpenMP.h:32-82, run with
 * one thread in raster order (the reference's in-place update is otherwise
 * racy when a stream cell carries the DEM nodata value). */
int or_stream_burn(float* z, const float* stream, int W, int H, int R, float nodata_dem,
                   float nodata_stream) {
    if (R < 0 || R > 5) return OR_ERR_ARG; /* MAX_BEVEL_RADIUS, StreamBurningCommon.h:17 */
    std::vector<Off> offs = circular_offsets(R);
    for (int sy = 0; sy < H; ++sy)
        for (int sx = 0; sx < W; ++sx) {
            size_t index = (size_t)sy * W + sx;
            if (stream[index] != nodata_stream) z[index] = 0.0f;
            for (const Off& p : offs) {
                int nx = sx + p.x, ny = sy + p.y;
                if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                size_t ni = (size_t)ny * W + nx;
                if (stream[ni] != nodata_stream && z[ni] != nodata_dem) {
                    float xd = (float)(sx - nx), yd = (float)(sy - ny);
                    float dist = std::sqrt(xd * xd + yd * yd);
                    z[index] = z[index] * (dist / (float)R);
                    break;
                }
            }
        }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Pit filling: PitFillingAlgorithmCPUOpenMP.h:68-242 (iter == 0 branch).   */
/* ------------------------------------------------------------------------ */

#define OR_PREC 1000

/* toIndex / toElev, PitFillingAlgorithmCPUOpenMP.h:55-65 */
static inline size_t to_index(float elev, float min_elev) {
    return (size_t)((float)OR_PREC * (elev - min_elev));
}
static inline float to_elev(size_t index, float min_elev) {
    return ((float)index) / (float)OR_PREC + min_elev;
}

/* Quantisation, PitFillingAlgorithmCPUOpenMP.h:84-106, plus the input
 * domain the reference handles deterministically: non-negative elevations
 * (the size_t conversion of a negative float is undefined) and a bucket
 * range whose float arithmetic stays exact (fl(fl(max-min)*1000) < 2^24,
 * otherwise the bucket array is indexed past its end). */
static int quantize_and_range(float* dem, float* spill, size_t n, float nodata,
                              float* min_out, float* max_out) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (size_t i = 0; i < n; ++i) {
        float elev = dem[i];
        if (elev == nodata) continue;
        if (elev < 0.0f || !(elev == elev)) {
            bad = 1;
            continue;
        }
        size_t q = to_index(elev, 0.0f);
        if (elev != 0.0 && q == 0) q = 1;
        dem[i] = to_elev(q, 0.0f);
    }
    if (bad) return OR_ERR_RANGE;
    std::memcpy(spill, dem, n * sizeof(float));
    /* thrust::minmax_element, PitFillingAlgorithmCPUOpenMP.h:112-116 */
    float mn = spill[0], mx = spill[0];
    for (size_t i = 1; i < n; ++i) {
        if (spill[i] < mn) mn = spill[i];
        if (mx < spill[i]) mx = spill[i];
    }
    float span = (mx - mn) * (float)OR_PREC;
    if (!(span < 16777216.0f)) return OR_ERR_RANGE;
    *min_out = mn;
    *max_out = mx;
    return OR_OK;
}

struct Cell {
    unsigned short x, y; /* PitFillingAlgorithmCPUOpenMP.h:43-52 */
};

/* LockQueue, LockQueue.h:23-75: a FIFO whose every operation is an OpenMP
 * critical section.  Kept because it shapes the CPU cost; the deque is
 * created on first push so an empty bucket costs one pointer. */
class LockQueue {
public:
    void push(const Cell& c) {
#pragma omp critical(or_lockqueue)
        {
            if (!q_) q_.reset(new std::queue<Cell>());
            q_->push(c);
        }
    }
    bool empty() {
        bool r;
#pragma omp critical(or_lockqueue)
        { r = !q_ || q_->empty(); }
        return r;
    }
    Cell front() {
        Cell c;
#pragma omp critical(or_lockqueue)
        { c = q_->front(); }
        return c;
    }
    void pop() {
#pragma omp critical(or_lockqueue)
        { q_->pop(); }
    }

private:
    std::unique_ptr<std::queue<Cell>> q_;
};

int or_pit_fill(float* spill, float* dem, int W, int H, float nodata) {
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535) return OR_ERR_ARG;
    const size_t n = (size_t)W * H;
    float minElev, maxElev;
    int st = quantize_and_range(dem, spill, n, nodata, &minElev, &maxElev);
    if (st != OR_OK) return st;

    size_t arraySize = (size_t)((maxElev - minElev) * OR_PREC + 1);
    std::vector<LockQueue> Q(arraySize);
    std::vector<char> processed(n, 0);

    /* seeds, PitFillingAlgorithmCPUOpenMP.h:121-144 */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t index = (size_t)y * W + x;
            if (spill[index] == nodata) continue;
            if (x == 0 || x == W - 1 || y == 0 || y == H - 1 || spill[index] == 0.0f) {
                size_t qIndex = to_index(spill[index], minElev);
                Q[qIndex].push(Cell{(unsigned short)x, (unsigned short)y});
                processed[index] = 1;
            }
        }

    /* flood, PitFillingAlgorithmCPUOpenMP.h:162-237 */
    static const int ndx[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
    static const int ndy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    for (size_t z = 0; z < arraySize; ++z) {
        float cSpill = to_elev(z, minElev);
        while (!Q[z].empty()) {
            Cell c = Q[z].front();
            Q[z].pop();
            int cx = c.x, cy = c.y;
            if (cSpill != spill[(size_t)cy * W + cx]) continue; /* dead-cell rule :178-181 */
            for (int k = 0; k < 8; ++k) {
                int nx = cx + ndx[k], ny = cy + ndy[k];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                size_t nIndex = (size_t)ny * W + nx;
                float nSpill = spill[nIndex];
                if (nSpill == nodata) continue;
                if (processed[nIndex]) continue;
                if (nSpill < cSpill) spill[nIndex] = cSpill;
                int nZOrig = (int)to_index(spill[nIndex], minElev);
                Q[nZOrig].push(Cell{(unsigned short)nx, (unsigned short)ny});
                processed[nIndex] = 1;
            }
        }
    }
    return OR_OK;
}

/* Second restatement: Dijkstra over values.  A reached cell takes
 * max(z, value of the first relaying neighbour); it relays only when its
 * value v satisfies toElev(toIndex(v)) == v.  Seeds are the edge and zero
 * cells; unreached cells keep the quantized elevation. */
int or_pit_fill_dijkstra(float* spill, float* dem, int W, int H, float nodata) {
    if (W <= 0 || H <= 0) return OR_ERR_ARG;
    const size_t n = (size_t)W * H;
    float minElev, maxElev;
    int st = quantize_and_range(dem, spill, n, nodata, &minElev, &maxElev);
    if (st != OR_OK) return st;
    typedef std::pair<float, size_t> Item;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    std::vector<char> done(n, 0);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            if (dem[i] == nodata) continue;
            if (x == 0 || y == 0 || x == W - 1 || y == H - 1 || dem[i] == 0.0f) {
                done[i] = 1;
                pq.push(Item(dem[i], i));
            }
        }
    while (!pq.empty()) {
        Item it = pq.top();
        pq.pop();
        float v = it.first;
        if (to_elev(to_index(v, minElev), minElev) != v) continue;
        int cx = (int)(it.second % W), cy = (int)(it.second / W);
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy) continue;
                int nx = cx + dx, ny = cy + dy;
                if (nx < 0 || ny < 0 || nx >= W || ny >= H) continue;
                size_t ni = (size_t)ny * W + nx;
                if (done[ni] || dem[ni] == nodata) continue;
                done[ni] = 1;
                float w = dem[ni] < v ? v : dem[ni];
                spill[ni] = w;
                pq.push(Item(w, ni));
            }
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Flow routing: FlowRoutingAlgorithmCPUOpenMP.h:53-249.                     */
/* ------------------------------------------------------------------------ */

/* neighbour order N, NE, E, SE, S, SW, W, NW (:105-114 / :186-195) */
static const int FR_DX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
static const int FR_DY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};

void or_flow_routing(int8_t* flowdir, int32_t* flatdist_out, const float* z, int W, int H) {
    const size_t n = (size_t)W * H;
    std::vector<int32_t> flatDist(n, 0);
    std::memset(flowdir, 0, 2 * n); /* clearData, :60-61 */
    std::vector<int> flats;

    /* performFlowRouting, :53-151 */
#pragma omp parallel
    {
        std::vector<int> local;
#pragma omp for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                int dxi = 0, dyi = 0;
                if (x == 0) dxi = -1;
                else if (x == W - 1) dxi = 1;
                if (y == 0) dyi = -1;
                else if (y == H - 1) dyi = 1;
                size_t index = (size_t)y * W + x;
                if (dxi != 0 || dyi != 0) {
                    flowdir[2 * index] = (int8_t)dxi;
                    flowdir[2 * index + 1] = (int8_t)dyi;
                    flatDist[index] = -1;
                    continue;
                }
                if (z[index] == 0.0f) { /* STREAM, DemData.h:15 */
                    flatDist[index] = -2;
                    continue;
                }
                float steepest = 0.0f;
                int sdx = 0, sdy = 0;
                for (int k = 0; k < 8; ++k) {
                    int nx = x + FR_DX[k], ny = y + FR_DY[k];
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                    float s = z[index] - z[(size_t)ny * W + nx];
                    if (s > 0.0f) {
                        if (k % 2 == 1) s = s / 1.41421356237309505f; /* MATH_SQRTTWO */
                        if (s > steepest) {
                            steepest = s;
                            sdx = FR_DX[k];
                            sdy = FR_DY[k];
                        }
                    }
                }
                if (steepest > 0.0f) {
                    flowdir[2 * index] = (int8_t)sdx;
                    flowdir[2 * index + 1] = (int8_t)sdy;
                    flatDist[index] = -1;
                } else {
                    local.push_back((int)index);
                }
            }
#pragma omp critical(or_flatlist)
        flats.insert(flats.end(), local.begin(), local.end());
    }

    /* performFlowRoutingOnFlatSurfaces, :153-249: rounds -1, 1, 2, ...
     * A value written in round k is k+1 (never k), so reading a neighbour
     * while it is being assigned cannot change a match. */
    int idToLookFor = -1;
    bool moreToDo = true;
    std::vector<char> assigned;
    while (moreToDo) {
        moreToDo = false;
        const int m = (int)flats.size();
        assigned.assign(m, 0);
        int any = 0;
#pragma omp parallel for schedule(static) reduction(| : any)
        for (int i = 0; i < m; ++i) {
            size_t index = (size_t)flats[i];
            int cx = (int)(index % W), cy = (int)(index / W);
            for (int k = 0; k < 8; ++k) {
                int nx = cx + FR_DX[k], ny = cy + FR_DY[k];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                size_t ni = (size_t)ny * W + nx;
                if (flatDist[ni] == idToLookFor && z[ni] <= z[index]) {
                    flatDist[index] = (idToLookFor == -1) ? 1 : idToLookFor + 1;
                    flowdir[2 * index] = (int8_t)FR_DX[k];
                    flowdir[2 * index + 1] = (int8_t)FR_DY[k];
                    assigned[i] = 1;
                    any = 1;
                }
            }
        }
        moreToDo = any != 0;
        std::vector<int> rest;
        rest.reserve(m);
        for (int i = 0; i < m; ++i)
            if (!assigned[i]) rest.push_back(flats[i]);
        flats.swap(rest);
        ++idToLookFor;
        if (idToLookFor == 0) idToLookFor = 1;
    }
    if (flatdist_out) std::memcpy(flatdist_out, flatDist.data(), n * sizeof(int32_t));
}

/* ------------------------------------------------------------------------ */
/* Flow tracing: FlowTracingAlgorithmCPUOpenMP.h:34-134.                     */
/* ------------------------------------------------------------------------ */

int64_t or_flow_tracing(int32_t* L, const int8_t* fd, const float* stream, int W, int H) {
    const size_t n = (size_t)W * H;
    std::memset(L, 0, n * sizeof(int32_t)); /* clearData, :48 */
    int64_t guarded = 0;
    std::vector<size_t> path;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int curX = x, curY = y;
            size_t index = (size_t)y * W + x;
            if (stream[index] >= 0.0) {
                L[index] = OR_FLOWTRACE_LAKE;
                continue;
            }
            if (L[index] < 0) continue;
            int iter = 0;
            while (true) {
                if (iter > 100000) { /* "Infinite loop?", :68-75 */
                    ++guarded;
                    break;
                }
                path.push_back(index);
                curX += fd[2 * index];
                curY += fd[2 * index + 1];
                index = (size_t)((long)curY * W + curX);
                if (curX < 0 || curX >= W || curY < 0 || curY >= H) {
                    for (size_t p : path) L[p] = OR_FLOWTRACE_OUTSIDE;
                    break;
                }
                if (stream[index] >= 0.0) {
                    for (size_t p : path) L[p] = (int32_t)(-stream[index]);
                    break;
                }
                if (L[index] < 0) {
                    int32_t v = L[index];
                    for (size_t p : path) L[p] = v;
                    break;
                }
                ++iter;
            }
            path.clear();
        }
    /* performConnectFlowPath, :115-134 */
    for (size_t i = 0; i < n; ++i) {
        if (L[i] >= 0) {
            int32_t ni = L[i];
            if (L[ni] < 0) L[i] = L[ni];
        }
    }
    return guarded;
}

/* ------------------------------------------------------------------------ */
/* Catchment border.                                                         */
/* ------------------------------------------------------------------------ */

void or_catchment_border(float* mc, const int32_t* L, int W, int H) {
    const size_t n = (size_t)W * H;
    std::vector<float> border(n, 0.0f); /* CatchmentBorderClass.h:66-74 clearData */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t index = (size_t)y * W + x;
            int32_t c = L[index];
            if (c == OR_FLOWTRACE_LAKE) continue;
            for (int k = 0; k < 8; ++k) {
                int nx = x + FR_DX[k], ny = y + FR_DY[k];
                if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
                int32_t v = L[(size_t)ny * W + nx];
                if (c < v && v != OR_FLOWTRACE_LAKE) {
                    border[index] = 1.0f;
                    break;
                }
            }
        }
    /* AddSurfaces with noDataValue -1, CatchmentBorderAlgorithm.h:60-69 */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        float d = mc[i];
        if (d != -1.0f) mc[i] = d + border[i];
    }
}

/* ------------------------------------------------------------------------ */
/* One realization, main.cpp:180-283.                                        */
/* ------------------------------------------------------------------------ */

int or_realization_from_noise(float* noise, const float* dem, const float* stream, int W,
                              int H, const float* filter, int radius, int model,
                              float nodata_dem, float nodata_stream, int bevel_radius,
                              float* mc, float* out_filled, int8_t* out_flowdir,
                              int32_t* out_labels) {
    const size_t n = (size_t)W * H;
    int st = or_convolve(noise, W, H, filter, radius, model);
    if (st) return st;
    /* rDem aliases rand, main.cpp:200-204 */
    or_realistic_dem(dem, noise, (int64_t)n, nodata_dem);
    st = or_stream_burn(noise, stream, W, H, bevel_radius, nodata_dem, nodata_stream);
    if (st) return st;
    std::vector<float> filled(n);
    st = or_pit_fill(filled.data(), noise, W, H, nodata_dem);
    if (st) return st;
    std::vector<int8_t> fd(2 * n);
    or_flow_routing(fd.data(), nullptr, filled.data(), W, H);
    std::vector<int32_t> labels(n);
    or_flow_tracing(labels.data(), fd.data(), stream, W, H);
    if (mc) or_catchment_border(mc, labels.data(), W, H);
    if (out_filled) std::memcpy(out_filled, filled.data(), n * sizeof(float));
    if (out_flowdir) std::memcpy(out_flowdir, fd.data(), 2 * n);
    if (out_labels) std::memcpy(out_labels, labels.data(), n * sizeof(int32_t));
    return OR_OK;
}

int or_realization(const float* dem, const float* stream, int W, int H, const float* filter,
                   int radius, int model, float mean, float std_dev, uint64_t seed,
                   uint32_t realization, float nodata_dem, float nodata_stream,
                   int bevel_radius, float* mc) {
    const size_t n = (size_t)W * H;
    std::vector<float> field(n);
    or_random_field(field.data(), (int64_t)n, seed, realization, 0u, mean, std_dev);
    return or_realization_from_noise(field.data(), dem, stream, W, H, filter, radius, model,
                                     nodata_dem, nodata_stream, bevel_radius, mc, nullptr,
                                     nullptr, nullptr);
}
