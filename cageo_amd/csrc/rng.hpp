// Counter-based normal generator used by the random-field stage and by the
// synthetic DEM generator.  Compiled for both host and gfx950 device code.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) gives four 32-bit words
// per counter; each pair becomes two N(0,1) values by Box-Muller evaluated in
// double precision with only correctly rounded IEEE operations (+, -, *, /,
// sqrt, floor), so host and device produce the same bits.
//
//   key     = (seed low 32 bits, seed high 32 bits)
//   counter = (block low 32 bits, realization, stream id, block high 32 bits)
//   cell i  <- block i/4, word i%4
//   u1 = ((w0 >> 8) + 1) / 2^24  in (0, 1],  u2 = (w1 >> 8) / 2^24  in [0, 1)
//   (n0, n1) = sqrt(-2 ln u1) * (cos 2 pi u2, sin 2 pi u2)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cageo {

struct U32x4 {
    uint32_t v[4];
};

__host__ __device__ inline uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *lo = (uint32_t)p;
    return (uint32_t)(p >> 32);
}

__host__ __device__ inline U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                               uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0, lo1;
        uint32_t hi0 = mulhilo32(0xD2511F53u, c0, &lo0);
        uint32_t hi1 = mulhilo32(0xCD9E8D57u, c2, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    U32x4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// ln(u), u in (0, 1] and normal: u = m 2^e, m in (sqrt(1/2), sqrt(2)],
// ln m = 2 atanh(s) with s = (m-1)/(m+1) and the atanh Taylor series to s^21.
__host__ __device__ inline double rng_log(double u) {
    uint64_t bits = (uint64_t)__builtin_bit_cast(uint64_t, u);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = __builtin_bit_cast(double, (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
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

// cos(2 pi t), sin(2 pi t), t in [0, 1): exact reduction to a quarter turn
// r in [-1/2, 1/2], a = r pi / 2, Taylor series of sin (to a^17) and cos
// (to a^18); |a| <= pi/4 keeps the truncation below 1e-19.
__host__ __device__ inline void rng_cossin_2pi(double t, double* c, double* s) {
    double x = 4.0 * t;
    double q = floor(x + 0.5);
    double r = x - q;
    double a = r * 1.5707963267948966;
    double a2 = a * a;
    double ps = -1.0 / 355687428096000.0;
    ps = ps * a2 + 1.0 / 1307674368000.0;
    ps = ps * a2 - 1.0 / 6227020800.0;
    ps = ps * a2 + 1.0 / 39916800.0;
    ps = ps * a2 - 1.0 / 362880.0;
    ps = ps * a2 + 1.0 / 5040.0;
    ps = ps * a2 - 1.0 / 120.0;
    ps = ps * a2 + 1.0 / 6.0;
    double sin_a = a - a * (a2 * ps);
    double pc = 1.0 / 6402373705728000.0;
    pc = pc * a2 - 1.0 / 20922789888000.0;
    pc = pc * a2 + 1.0 / 87178291200.0;
    pc = pc * a2 - 1.0 / 479001600.0;
    pc = pc * a2 + 1.0 / 3628800.0;
    pc = pc * a2 - 1.0 / 40320.0;
    pc = pc * a2 + 1.0 / 720.0;
    pc = pc * a2 - 1.0 / 24.0;
    pc = pc * a2 + 0.5;
    double cos_a = 1.0 - a2 * pc;
    int qi = ((int)q) & 3;
    if (qi == 0) {
        *c = cos_a;
        *s = sin_a;
    } else if (qi == 1) {
        *c = -sin_a;
        *s = cos_a;
    } else if (qi == 2) {
        *c = -cos_a;
        *s = -sin_a;
    } else {
        *c = sin_a;
        *s = -cos_a;
    }
}

__host__ __device__ inline void box_muller(uint32_t a, uint32_t b, double* n0, double* n1) {
    double u1 = (double)((a >> 8) + 1u) * (1.0 / 16777216.0);
    double u2 = (double)(b >> 8) * (1.0 / 16777216.0);
    double rad = sqrt(-2.0 * rng_log(u1));
    double c, s;
    rng_cossin_2pi(u2, &c, &s);
    *n0 = rad * c;
    *n1 = rad * s;
}

// Four N(0,1) draws for block `b` (cells 4b .. 4b+3).
__host__ __device__ inline void normal4(uint64_t b, uint32_t realization, uint32_t stream_id,
                                        uint64_t seed, float out[4]) {
    U32x4 x = philox4x32_10((uint32_t)b, realization, stream_id, (uint32_t)(b >> 32),
                            (uint32_t)seed, (uint32_t)(seed >> 32));
    double g0, g1, g2, g3;
    box_muller(x.v[0], x.v[1], &g0, &g1);
    box_muller(x.v[2], x.v[3], &g2, &g3);
    out[0] = (float)g0;
    out[1] = (float)g1;
    out[2] = (float)g2;
    out[3] = (float)g3;
}

}  // namespace cageo
