// Process-convolution filter construction (host, once per run).
//
// ConvolutionAlgorithm<T>::createFilter, ConvolutionAlgorithm.cpp:18-45:
//   1. correlogram C on the (2r+1)^2 lag grid (createCorrelogram :49-82);
//      exponential exp(-d / ((PR/3)/cell)) in float, Gaussian
//      exp(-d^2 / ((PR/sqrt 3)/cell)^2) in double;
//   2. K = shift_{L/2}( IDFT( |DFT(C)|^(1/2) ) ) in double
//      (createFilterFromCorrelogram :84-148, FFTW r2c/c2r there);
//   3. Gaussian only: the separable vector u0 * sqrt(s0) of the leading
//      singular pair of the float K (createKernelVector :150-172, Eigen's
//      JacobiSVD there).  The reference leaves the sign of u0 to Eigen; this
//      library takes the centre tap positive.
// The DFTs here are direct row-column sums with exact twiddle indexing.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "cageo_hip.h"

namespace {

typedef std::complex<double> cd;

// In-place 1-D DFT of `len` samples with stride `stride`; sign -1 forward.
void dft_1d(cd* data, int len, int stride, int sign, const std::vector<cd>& roots) {
    std::vector<cd> tmp(len);
    for (int k = 0; k < len; ++k) {
        cd acc(0.0, 0.0);
        for (int j = 0; j < len; ++j) {
            cd w = roots[(size_t)((long)k * j % len)];
            if (sign > 0) w = std::conj(w);
            acc += data[(size_t)j * stride] * w;
        }
        tmp[k] = acc;
    }
    for (int k = 0; k < len; ++k) data[(size_t)k * stride] = tmp[k];
}

void dft_2d(std::vector<cd>& a, int L, int sign) {
    std::vector<cd> roots(L);
    for (int k = 0; k < L; ++k) {
        double ang = -2.0 * M_PI * (double)k / (double)L;
        roots[k] = cd(std::cos(ang), std::sin(ang));
    }
    for (int y = 0; y < L; ++y) dft_1d(&a[(size_t)y * L], L, 1, sign, roots);
    for (int x = 0; x < L; ++x) dft_1d(&a[x], L, L, sign, roots);
}

}  // namespace

extern "C" int cageo_filter_radius(float practical_range, float cell_size) {
    if (!(cell_size > 0.0f) || !(practical_range >= 0.0f)) return CAGEO_ERR_ARG;
    return (int)std::ceil((practical_range / cell_size) * 2.5f);
}

extern "C" int cageo_create_filter(int model, float pr, float cell, int radius, float* out) {
    if (!out || radius < 0 || radius > 4096 || !(cell > 0.0f) ||
        (model != CAGEO_FILTER_EXPONENTIAL && model != CAGEO_FILTER_GAUSSIAN))
        return CAGEO_ERR_ARG;
    const int L = 2 * radius + 1;
    const size_t LL = (size_t)L * L;

    std::vector<cd> a(LL);
    for (int y = -radius, i = 0; y <= radius; ++y)
        for (int x = -radius; x <= radius; ++x, ++i) {
            float c;
            if (model == CAGEO_FILTER_EXPONENTIAL) {
                float dist = std::sqrt((float)(x * x + y * y));
                c = (float)std::exp(-dist / ((pr / 3.0f) / cell));
            } else {
                double range = (double)((pr / 1.73205080757f) / cell);
                c = (float)std::exp(-(double)(x * x + y * y) / std::pow(range, 2));
            }
            a[i] = cd((double)c, 0.0);
        }

    dft_2d(a, L, -1);
    for (cd& v : a) v = cd(std::sqrt(std::sqrt(std::norm(v))), 0.0);
    dft_2d(a, L, +1);

    std::vector<float> k2(LL);
    const int s = L / 2;
    for (int j = 0; j < L; ++j)
        for (int i = 0; i < L; ++i)
            k2[(size_t)((j + s) % L) * L + (i + s) % L] = (float)(a[(size_t)j * L + i].real() / (double)LL);

    if (model == CAGEO_FILTER_EXPONENTIAL) {
        for (size_t i = 0; i < LL; ++i) out[i] = k2[i];
        return CAGEO_OK;
    }

    // Leading singular pair of the float kernel by subspace (power) iteration
    // on K^T K; the kernel is numerically rank one so this converges in a
    // handful of steps.
    std::vector<double> v(L, 1.0), u(L), w(L);
    for (int it = 0; it < 64; ++it) {
        for (int j = 0; j < L; ++j) {
            double acc = 0.0;
            for (int i = 0; i < L; ++i) acc += (double)k2[(size_t)j * L + i] * v[i];
            u[j] = acc;
        }
        double nrm = 0.0;
        for (int i = 0; i < L; ++i) {
            double acc = 0.0;
            for (int j = 0; j < L; ++j) acc += (double)k2[(size_t)j * L + i] * u[j];
            w[i] = acc;
            nrm += acc * acc;
        }
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0)) return CAGEO_ERR_RANGE;
        for (int i = 0; i < L; ++i) v[i] = w[i] / nrm;
    }
    double s0sq = 0.0;
    for (int j = 0; j < L; ++j) {
        double acc = 0.0;
        for (int i = 0; i < L; ++i) acc += (double)k2[(size_t)j * L + i] * v[i];
        u[j] = acc;
        s0sq += acc * acc;
    }
    const double s0 = std::sqrt(s0sq);
    const double sign = u[radius] < 0.0 ? -1.0 : 1.0;
    for (int j = 0; j < L; ++j) out[j] = (float)(sign * (u[j] / s0) * std::sqrt(s0));
    return CAGEO_OK;
}

// Low-rank separable form of a symmetric 2-D filter (no reference
// counterpart: a fast mode of the exponential convolution).  For f = f^T
// (the 2-D filters are symmetric under transposition, to rounding),
//   f = sum_k lambda_k v_k v_k^T   (eigen-decomposition, cyclic Jacobi in
// double), so the 2-D convolution is a sum of separable ones.  Terms are
// kept in decreasing |lambda| until the Frobenius norm of the remainder is
// at most `tol` times that of f, at most `kmax` of them.
extern "C" int cageo_filter_low_rank(const float* taps, int radius, int kmax, double tol,
                                     float* vectors, float* values, int* k_out,
                                     double* residual_out) {
    if (!taps || radius < 0 || kmax < 1 || !vectors || !values || !k_out) return CAGEO_ERR_ARG;
    const int n = 2 * radius + 1;
    std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
    // The filters are symmetric up to the rounding of their construction:
    // the symmetric part is decomposed, an asymmetry above 1e-6 of the
    // largest tap is refused.
    double fmax = 0.0, asym = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double fij = taps[(size_t)i * n + j], fji = taps[(size_t)j * n + i];
            fmax = std::max(fmax, std::fabs(fij));
            asym = std::max(asym, std::fabs(fij - fji));
            a[(size_t)i * n + j] = 0.5 * (fij + fji);
        }
    if (asym > 1e-6 * fmax) return CAGEO_ERR_ARG;
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    double fro2 = 0.0;
    for (double x : a) fro2 += x * x;
    // cyclic Jacobi: rotate away each off-diagonal pair until negligible
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
        if (off <= 1e-30 * fro2) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {  // eigenvectors (columns of v)
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return std::fabs(a[(size_t)x * n + x]) > std::fabs(a[(size_t)y * n + y]);
    });
    // remainder norm^2 after K terms = sum of the dropped eigenvalues^2
    double rest = 0.0;
    for (int i = 0; i < n; ++i) rest += a[(size_t)i * n + i] * a[(size_t)i * n + i];
    int k = 0;
    while (k < kmax && k < n && rest > tol * tol * fro2) {
        const int e = order[k];
        const double lam = a[(size_t)e * n + e];
        rest -= lam * lam;
        values[k] = (float)lam;
        for (int i = 0; i < n; ++i) vectors[(size_t)k * n + i] = (float)v[(size_t)i * n + e];
        ++k;
    }
    *k_out = k;
    if (residual_out) *residual_out = std::sqrt(std::max(rest, 0.0) / fro2);
    return rest <= tol * tol * fro2 ? CAGEO_OK : CAGEO_ERR_RANGE;
}
