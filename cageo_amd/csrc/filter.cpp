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
