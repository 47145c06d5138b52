// Seeded synthetic DEM and stream network (host).  The reference ships a
// 1560x1341 sample DEM whose payload is missing, so runs at any size use
// this generator instead:
//   DEM  = 100 m + 0.005*cell*x + 40 Gaussian hills + N(0, 0.3^2)
//   streams: D8 flow accumulation > threshold on the 5x5 box-smoothed DEM,
//            filled with epsilon steps (the result of Barnes, Lehman and
//            Mulla 2014, "Priority-flood: an optimal depression-filling and
//            watershed-labeling algorithm", +epsilon variant, computed by the
//            relaxation of Planchon and Darboux 2001, "A fast, simple and
//            versatile algorithm to fill the depressions of digital elevation
//            models"), one float id >= 1 per stream link (a new id at every
//            channel head and confluence, in decreasing elevation order),
//            nodata elsewhere.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "cageo_hip.h"
#include "rng.hpp"

namespace {

// Uniform double in [0, 1) from Philox word pairs (53 bits).
struct Uniforms {
    uint64_t seed;
    uint32_t stream;
    uint64_t ctr = 0;
    uint32_t buf[4];
    int left = 0;
    uint32_t next32() {
        if (left == 0) {
            cageo::U32x4 x = cageo::philox4x32_10((uint32_t)ctr, 0u, stream, (uint32_t)(ctr >> 32),
                                                  (uint32_t)seed, (uint32_t)(seed >> 32));
            ++ctr;
            std::memcpy(buf, x.v, sizeof(buf));
            left = 4;
        }
        return buf[4 - left--];
    }
    double uniform() {
        uint64_t a = next32() >> 5, b = next32() >> 6;
        return (double)(a * 67108864ull + b) * (1.0 / 9007199254740992.0);
    }
};

// Static row partition over the host's hardware threads.
template <class F>
void parallel_rows(int rows, F body) {
    int nt = (int)std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    nt = std::min(nt, rows);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=, &body]() {
            for (int r = (int)((int64_t)rows * t / nt); r < (int)((int64_t)rows * (t + 1) / nt); ++r)
                body(r);
        });
    for (auto& x : th) x.join();
}

// Row bands [r0, r1) of `rows` over the host's hardware threads.
template <class F>
void parallel_bands(int rows, F body) {
    int nt = (int)std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    nt = std::max(1, std::min(nt, rows / 8));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=, &body]() {
            body((int)((int64_t)rows * t / nt), (int)((int64_t)rows * (t + 1) / nt));
        });
    for (auto& x : th) x.join();
}

// Relaxed atomic access to a float shared between the band threads.
inline float load_relaxed(const float* p) {
    const uint32_t b = __atomic_load_n(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED);
    float f;
    std::memcpy(&f, &b, sizeof(f));
    return f;
}
inline void store_relaxed(float* p, float v) {
    uint32_t b;
    std::memcpy(&b, &v, sizeof(b));
    __atomic_store_n(reinterpret_cast<uint32_t*>(p), b, __ATOMIC_RELAXED);
}

const int DX8[8] = {0, 1, 1, 1, 0, -1, -1, -1};
const int DY8[8] = {-1, -1, 0, 1, 1, 1, 0, -1};

}  // namespace

extern "C" int cageo_synth_dem(int W, int H, uint64_t seed, float cell, int acc_threshold,
                               float nodata_stream, float* dem, float* stream) {
    if (W < 3 || H < 3 || !dem || !stream || !(cell > 0.0f)) return CAGEO_ERR_ARG;
    const size_t n = (size_t)W * H;

    // Hills.
    Uniforms u{seed, 0x51A7E000u};
    struct Hill {
        double cx, cy, h, s;
    };
    std::vector<Hill> hills(40);
    for (Hill& hl : hills) {
        hl.h = 5.0 + 55.0 * u.uniform();
        hl.s = (50.0 + 350.0 * u.uniform()) / (double)cell;  // in cells
        hl.cx = W * u.uniform();
        hl.cy = H * u.uniform();
    }

    // Base surface + hills + white noise (normal stream 0x51A7E001).
    const double tilt = 0.005 * (double)cell;
    parallel_rows(H, [&](int y) {
        for (int x = 0; x < W; ++x) {
            double z = 100.0 + tilt * x;
            for (const Hill& hl : hills) {
                double dx = x - hl.cx, dy = y - hl.cy;
                double r2 = (dx * dx + dy * dy) / (2.0 * hl.s * hl.s);
                if (r2 < 40.0) z += hl.h * std::exp(-r2);
            }
            dem[(size_t)y * W + x] = (float)z;
        }
    });
    const int64_t nb = (int64_t)((n + 3) / 4);
    const int chunks = (int)std::min<int64_t>(nb, 4096);
    parallel_rows(chunks, [&](int c) {
        for (int64_t b = nb * c / chunks; b < nb * (c + 1) / chunks; ++b) {
            float g[4];
            cageo::normal4((uint64_t)b, 0u, 0x51A7E001u, seed, g);
            for (int k = 0; k < 4; ++k) {
                size_t i = 4 * (size_t)b + k;
                if (i < n) dem[i] = dem[i] + 0.3f * g[k];
            }
        }
    });

    // 5x5 box smoothing (mean over the in-grid part of the window).
    std::vector<float> sm(n);
    parallel_rows(H, [&](int y) {
        for (int x = 0; x < W; ++x) {
            double s = 0.0;
            int c = 0;
            for (int dy = -2; dy <= 2; ++dy) {
                int yy = y + dy;
                if (yy < 0 || yy >= H) continue;
                for (int dx = -2; dx <= 2; ++dx) {
                    int xx = x + dx;
                    if (xx < 0 || xx >= W) continue;
                    s += dem[(size_t)yy * W + xx];
                    ++c;
                }
            }
            sm[(size_t)y * W + x] = (float)(s / c);
        }
    });

    // Depression filling with epsilon steps from the grid edge: every interior
    // cell ends strictly above a neighbour.  The result is that of the
    // priority flood with epsilon steps (Barnes, Lehman and Mulla 2014, the
    // +epsilon variant: cells leave a priority queue in (elevation, index)
    // order and a cell takes max(z, nextafter(z of the neighbour that reached
    // it))), which is the greatest fixed point of
    //   F(n) = max(sm(n), nextafter(min over the 8 neighbours c of F(c)))
    // with the edge cells at sm: the neighbour that reaches n first is its
    // lowest one, and by induction over the queue order no fixed point lies
    // above it.  That fixed point is computed here by the Planchon and
    // Darboux (2001) relaxation from +infinity: sweeps in four raster
    // orders, each over bands of rows in parallel, until a round of sweeps
    // changes nothing.  Values only decrease and the map is monotone, so the
    // order of the updates (and reads of a neighbour band's rows while it
    // sweeps) does not change the limit.
    {
        std::vector<float> w(n, INFINITY);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                if (x == 0 || y == 0 || x == W - 1 || y == H - 1) w[(size_t)y * W + x] = sm[(size_t)y * W + x];
        for (;;) {
            std::atomic<long long> changed{0};
            for (int dir = 0; dir < 4; ++dir) {
                const bool down = dir < 2, right = (dir & 1) == 0;
                parallel_bands(H - 2, [&](int r0, int r1) {
                    long long ch = 0;
                    for (int k = 0; k < r1 - r0; ++k) {
                        const int y = 1 + (down ? r0 + k : r1 - 1 - k);
                        for (int j = 0; j < W - 2; ++j) {
                            const int x = right ? 1 + j : W - 2 - j;
                            const size_t i = (size_t)y * W + x;
                            const float cur = load_relaxed(&w[i]);
                            if (cur == sm[i]) continue;  // cannot go lower
                            float m = load_relaxed(&w[i - W - 1]);
                            m = std::min(m, load_relaxed(&w[i - W]));
                            m = std::min(m, load_relaxed(&w[i - W + 1]));
                            m = std::min(m, load_relaxed(&w[i - 1]));
                            m = std::min(m, load_relaxed(&w[i + 1]));
                            m = std::min(m, load_relaxed(&w[i + W - 1]));
                            m = std::min(m, load_relaxed(&w[i + W]));
                            m = std::min(m, load_relaxed(&w[i + W + 1]));
                            const float v = std::max(sm[i], std::nextafter(m, INFINITY));
                            if (v < cur) {
                                store_relaxed(&w[i], v);
                                ++ch;
                            }
                        }
                    }
                    changed += ch;
                });
            }
            if (changed.load() == 0) break;
        }
        sm.swap(w);
    }

    // D8 steepest descent on the filled surface; edge cells drain out.
    std::vector<int32_t> down(n, -1);
    parallel_rows(H - 2, [&](int yy) {
        const int y = yy + 1;
        for (int x = 1; x < W - 1; ++x) {
            size_t i = (size_t)y * W + x;
            double best = 0.0;
            int bk = -1;
            for (int k = 0; k < 8; ++k) {
                size_t ni = (size_t)(y + DY8[k]) * W + (x + DX8[k]);
                double s = (double)sm[i] - (double)sm[ni];
                if (k & 1) s /= std::sqrt(2.0);
                if (s > best) {
                    best = s;
                    bk = k;
                }
            }
            if (bk >= 0) down[i] = (int32_t)((y + DY8[bk]) * W + (x + DX8[bk]));
        }
    });

    // Flow accumulation: a cell passes its total downstream once all its
    // donors have (a topological order of the D8 tree; the sums do not depend
    // on which order).
    std::vector<uint32_t> acc(n, 1);
    {
        std::vector<uint8_t> donors(n, 0);
        for (size_t i = 0; i < n; ++i)
            if (down[i] >= 0) ++donors[down[i]];
        std::vector<uint32_t> ready;
        ready.reserve(1 << 16);
        for (size_t i = 0; i < n; ++i)
            if (donors[i] == 0) ready.push_back((uint32_t)i);
        while (!ready.empty()) {
            const uint32_t i = ready.back();
            ready.pop_back();
            const int32_t d = down[i];
            if (d < 0) continue;
            acc[d] += acc[i];
            if (--donors[d] == 0) ready.push_back((uint32_t)d);
        }
    }

    // Stream cells and link ids (upstream first): the stream cells in
    // decreasing elevation order, ties by index.
    std::vector<uint32_t> order;
    for (size_t i = 0; i < n; ++i)
        if (acc[i] > (uint32_t)acc_threshold) order.push_back((uint32_t)i);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return sm[a] > sm[b] || (sm[a] == sm[b] && a < b);
    });
    std::vector<uint8_t> nup(n, 0);
    for (size_t i = 0; i < n; ++i)
        if (acc[i] > (uint32_t)acc_threshold && down[i] >= 0 &&
            acc[down[i]] > (uint32_t)acc_threshold)
            nup[down[i]] = (uint8_t)std::min(255, nup[down[i]] + 1);
    std::vector<float> id(n, 0.0f);
    float next_id = 1.0f;
    for (uint32_t i : order) {
        if (nup[i] != 1) id[i] = next_id++;  // head or confluence
        if (down[i] >= 0 && acc[down[i]] > (uint32_t)acc_threshold && nup[down[i]] == 1)
            id[down[i]] = id[i];
    }
    for (size_t i = 0; i < n; ++i) stream[i] = acc[i] > (uint32_t)acc_threshold ? id[i] : nodata_stream;
    return CAGEO_OK;
}
