// Pit filling: PitFillingAlgorithm<T>::execute, PitFillingAlgorithm.h:88-98,
// parity target performPitFilling, PitFillingAlgorithmCPUOpenMP.h:68-242.
//
// What the reference computes.  After quantising to mm (:84-106) it floods
// from the seeds (non-nodata edge cells and cells at 0) through buckets of
// 1/1000 m in increasing order (:162-237).  A popped cell c expands its
// unprocessed 8-neighbours n to max(z_n, spill_c); every cell is pushed once,
// into bucket toIndex(v) of its value v, and when popped it expands only if
// toElev(toIndex(v)) == v (:178-181).  Values that fail that round trip
// ("dead" values, e.g. 0.251 m) never expand, so with P the result
//   P(n) = max(Z(n), min { P(c) : c an 8-neighbour, reached, live(P(c)) }),
// seeds fixed at Z, unreached cells keep Z.
//
// How it is computed here, without a global priority queue.  Cells are
// FIXED (value known to equal P; live or dead) or OPEN.  Seeds start FIXED.
// Each round:
//   fill:    monotone relaxation from +inf of W(n) = max(Z(n), min relay(c))
//            over the OPEN cells, relay(c) = W(c) for OPEN and live FIXED
//            cells.  Every relay the reference makes is available, so at
//            convergence W <= P on the OPEN cells.  During the relaxation an
//            OPEN cell with a live FIXED neighbour c, P(c) <= Z(n), is fixed
//            at Z(n) at once: P(n) <= max(Z(n), P(c)) = Z(n) <= P(n).  A cell
//            fixed dead stops relaying.
//   certify: after convergence an OPEN cell with a live FIXED neighbour c,
//            P(c) <= W(n), is fixed: P(n) <= max(Z(n), P(c)) <= W(n) <= P(n).
//   next:    OPEN cells restart at +inf.
// A round fixes at least the first OPEN cell the reference reaches (every
// relay path into it through OPEN cells has bottleneck >= its value), so a
// round that fixes nothing leaves only unreached cells; they keep Z.
//
// Execution: one persistent kernel runs every round.  Passes over a
// de-duplicated worklist of 64x64 tiles are separated by grid barriers; the
// first pass of a phase takes the tiles that were open after their first
// solve in the phase before (OPEN cells never reappear, so that covers every
// open tile; the fill's list is cut down to the tiles still open, since most
// tiles certified at all are certified after the first pass), and the
// fill's first pass is also the restart (OPEN cells, its own and its
// halo's, read as +inf; it rewrites its own);
// inside a tile, rows are solved exactly one at a time (wave scans along the
// row, rows revisited while a neighbouring row changed).  Every decision that
// ends a loop is read by all workgroups after a barrier, so they leave
// together.
//
// Every comparison is on floats; the only arithmetic is the quantisation and
// the round-trip test, both evaluated exactly as the reference does.
#include <vector>

#include "tiles.hpp"

namespace cageo {

namespace {

constexpr float PF_INF = __builtin_huge_valf();

struct PfShared {
    float2 rel[HALO_W * HALO_W];  // x: relay (OPEN or live FIXED), y: live FIXED only
    uint32_t rowdirty[TILE];      // strip edge rows marked by the neighbouring strips
    uint32_t wcol[TILE_THREADS / 64], ecol[TILE_THREADS / 64];  // changed rows of columns 0, TILE-1
    uint32_t bits;
    uint32_t any;
    uint32_t next;  // the workgroup's next tile index in the pass
    uint32_t steps; // row steps of the tile's solve, all waves
    uint32_t ringrows[2];  // resident rounds: tile rows next to changed halo cells
    uint32_t pfront, pback;  // entries this workgroup listed in the pass (thread 0's)
};

// Z of a thread's 16 cells (row k of its strip), held in registers by the
// persistent kernel: a vector, so that the sweeps' read of row k (the same k
// across the wave) is an indexed register move, where an array would live
// in scratch.  Without a Z tile in LDS (35.4 KB a workgroup) four workgroups
// share a CU.
typedef float ZRow __attribute__((ext_vector_type(STRIP)));

// The resident rounds keep the tile's Z in LDS (their kernel has no
// registers to spare).
struct PfSharedRes : PfShared {
    float zt[TILE * TILE];
};

// Batches of a tile's own rows in load_tile_own (each a round of global
// loads; the first carries the halo ring too).
constexpr int PF_LOAD_BATCHES = 1;

// run_passes result of a launch being abandoned at a barrier.
constexpr int PF_ABANDONED = -2;

// A solve of at least this many row steps (all waves; one sweep is 64)
// ranks its tile heavy for the next pass (see Ranked).
constexpr uint32_t PF_HEAVY = 256;

__device__ inline float min8x(const float2* s, int i) {
    float a = fminf(s[i - HALO_W - 1].x, s[i - HALO_W].x);
    float b = fminf(s[i - HALO_W + 1].x, s[i - 1].x);
    float c = fminf(s[i + 1].x, s[i + HALO_W - 1].x);
    float d = fminf(s[i + HALO_W].x, s[i + HALO_W + 1].x);
    return fminf(fminf(a, b), fminf(c, d));
}
// Cells of a row reached from the `seed` lanes by steps x-1 -> x allowed
// by `gate_up` and x+1 -> x allowed by `gate_down` (64-bit lane masks).
// Upward, R(x) = S(x) | (G(x) & R(x-1)) is the carry recurrence of a binary
// addition with generate bits S and propagate bits G & ~S, i.e. of
// (S | G) + S: R(x) is the carry into bit x+1, so R = carries >> 1 with the
// carry out of bit 63 as R(63).  Downward is the same on bit-reversed masks.
// A few scalar instructions instead of a doubling loop.
__device__ inline unsigned long long spread_up(unsigned long long s, unsigned long long g) {
    const unsigned long long x = s | g;
    const unsigned long long sum = x + s;
    // carry out of bit 63 from the operands' and the sum's top bits (scalar
    // bit operations, where an overflow test became a vector compare)
    const unsigned long long out = ((x & s) | ((x | s) & ~sum)) >> 63;
    return ((sum ^ x ^ s) >> 1) | (out << 63);
}
__device__ inline unsigned long long spread_row(unsigned long long seed, unsigned long long gate_up,
                                                unsigned long long gate_down) {
    const unsigned long long down = __builtin_bitreverse64(
        spread_up(__builtin_bitreverse64(seed), __builtin_bitreverse64(gate_down)));
    return spread_up(seed, gate_up) | down;
}

// Relays of the six cells in the rows above and below (the row neighbours
// enter through the row scan).
__device__ inline float min6x(const float2* s, int i) {
    float a = fminf(s[i - HALO_W - 1].x, s[i - HALO_W].x);
    float b = fminf(s[i - HALO_W + 1].x, s[i + HALO_W - 1].x);
    float c = fminf(s[i + HALO_W].x, s[i + HALO_W + 1].x);
    return fminf(fminf(a, b), c);
}
__device__ inline float min8y(const float2* s, int i) {
    float a = fminf(s[i - HALO_W - 1].y, s[i - HALO_W].y);
    float b = fminf(s[i - HALO_W + 1].y, s[i - 1].y);
    float c = fminf(s[i + 1].y, s[i + HALO_W - 1].y);
    float d = fminf(s[i + HALO_W].y, s[i + HALO_W + 1].y);
    return fminf(fminf(a, b), fminf(c, d));
}


__device__ inline uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One raster of a launch.
struct PfRaster {
    const float* zq;
    float* w;
    uint8_t* st;
    int W, H, ntx, nty;
    float nodata;
    const uint32_t* __restrict__ scal;  // min/max/bad of the quantized raster
    uint32_t* status;      // the raster's sticky status word
};

// A launch over one raster (`one`, rs null) or over the nr rasters of a
// table, whose 64x64 blocks are numbered globally (bmap).
struct PfArgs {
    PfRaster one;
    const PfRaster* __restrict__ rs;
    const int2* __restrict__ bmap;  // per block: (raster, block within it)
    uint2* rinfo;                   // per raster: minimum (float bits), domain ok
    int nr, total;
    int* work;            // 7 lists of `total` entries: 3 rotating pass lists, 3 phase lists,
                          // the compacted fill list
    uint32_t* mark;       // per-block epoch stamps
    uint32_t* topen;      // per-block: holds OPEN cells (written by every solve)
    uint32_t* tcost;      // per-block: row steps of its last solve (list ranking)
    uint32_t* rows;       // per-block dirty-row masks: 2 pass parities x 64 bits
    uint32_t* ctrl;       // control block (CTL_*)
    uint32_t* cert;       // per raster: 3 rotating per-round fixed-cell counters
    GridBar gb;           // grid barrier
    int max_passes;
    int resident_cap;     // co-resident workgroups of k_pf_resident (0: no resident rounds)
};

// The raster of global block g, g's number within it, and the raster's
// minimum and domain flag (`mn1`, `ok1` for a one-raster launch; the
// per-raster words of the table once the launch has written them).
struct PfBlock {
    PfRaster R;
    int r, lb, gbase;
    float mn;
    bool ok;
};
template <bool MULTI>
__device__ inline PfBlock pf_block(const PfArgs& a, int g, float mn1, bool ok1) {
    PfBlock b;
    g = __builtin_amdgcn_readfirstlane(g);
    if constexpr (!MULTI) {
        b.R = a.one;
        b.r = 0;
        b.lb = g;
        b.mn = mn1;
        b.ok = ok1;
    } else {
        const int2 m = a.bmap[g];
        b.r = __builtin_amdgcn_readfirstlane(m.x);
        b.lb = __builtin_amdgcn_readfirstlane(m.y);
        b.R = a.rs[b.r];
        const uint2 inf = a.rinfo[b.r];
        b.mn = __uint_as_float(__builtin_amdgcn_readfirstlane(inf.x));
        b.ok = __builtin_amdgcn_readfirstlane(inf.y) != 0u;
    }
    b.gbase = g - b.lb;
    return b;
}

// Domain of the reference's arithmetic: non-negative elevations and a bucket
// array fl(fl(max-min)*1000)+1 that float arithmetic indexes exactly (beyond
// 2^24 the reference reads past it).  Sets the raster's minimum.
// The scalars were written by an earlier launch and are read-only here, so
// plain (cacheable) loads are exact.
__device__ inline bool pf_raster_ok(const PfRaster& R, float& mn) {
    const uint32_t bad = R.scal[SC_BAD];
    mn = ordered_to_float(R.scal[SC_MIN]);
    const float mx = ordered_to_float(R.scal[SC_MAX]);
    const float span = (mx - mn) * 1000.0f;
    return !bad && span < 16777216.0f;
}

// Loads the tile's relays with a one-cell halo.  The state byte and the
// value of a cell are loaded together (no load waits on another), over a
// fully unrolled loop, so the loads of all iterations are in flight at once.
// With `restart`, OPEN cells relay nothing: their stored values are the
// last round's (a neighbour that lowers one in this pass activates the tile).
__device__ inline void load_tile(PfShared& S, const PfRaster& a, int x0, int y0, bool restart) {
    constexpr int N = HALO_W * HALO_W;
    constexpr int ITER = (N + TILE_THREADS - 1) / TILE_THREADS;
    constexpr int CHUNK = (ITER + 1) / 2;  // two batches keep the register use down
#pragma unroll
    for (int c0 = 0; c0 < ITER; c0 += CHUNK) {
        uint8_t sv[CHUNK];
        float wv[CHUNK];
#pragma unroll
        for (int j = 0; j < CHUNK; ++j) {
            const int k = threadIdx.x + (c0 + j) * TILE_THREADS;
            sv[j] = PF_NODATA;
            wv[j] = PF_INF;
            if (c0 + j < ITER && k < N) {
                const int hy = k / HALO_W, hx = k - hy * HALO_W;
                const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
                if (gx >= 0 && gy >= 0 && gx < a.W && gy < a.H) {
                    const size_t gi = (size_t)gy * a.W + gx;
                    sv[j] = a.st[gi];
                    wv[j] = a.w[gi];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < CHUNK; ++j) {
            const int k = threadIdx.x + (c0 + j) * TILE_THREADS;
            if (c0 + j < ITER && k < N) {
                const uint8_t st = sv[j];
                const float r = (st == PF_FIXED_LIVE || (st == PF_OPEN && !restart)) ? wv[j] : PF_INF;
                S.rel[k] = make_float2(r, st == PF_FIXED_LIVE ? r : PF_INF);
            }
        }
    }
    if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        S.bits = 0;
        S.any = 0;
        S.steps = 0;
    }
}

// load_tile and the tile's own states in fewer dependent rounds: each
// thread loads its strip's 16 cells (state, value and, with zq, the
// quantized elevation into zr) in PF_LOAD_BATCHES batches of rows, the
// first with the halo ring's cells: one round of global loads a batch.
// Returns the thread's OPEN cells (bit k = row k of its strip).
__device__ inline uint32_t load_tile_own(PfShared& S, const PfRaster& a, int x0, int y0,
                                        bool restart, const float* zq, ZRow& zr) {
    constexpr int RING = 4 * HALO_W - 4;  // 260 cells
    constexpr int RIT = (RING + TILE_THREADS - 1) / TILE_THREADS;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gx = x0 + lx;
    auto relay = [&](uint8_t st, float w) {
        const float r = (st == PF_FIXED_LIVE || (st == PF_OPEN && !restart)) ? w : PF_INF;
        return make_float2(r, st == PF_FIXED_LIVE ? r : PF_INF);
    };
    uint8_t rs[RIT];
    float rw[RIT];
    int rl[RIT];
#pragma unroll
    for (int j = 0; j < RIT; ++j) {
        const int k = threadIdx.x + j * TILE_THREADS;
        rs[j] = PF_NODATA;
        rw[j] = PF_INF;
        rl[j] = -1;
        if (k < RING) {
            int hx, hy;
            if (k < HALO_W) {
                hy = 0;
                hx = k;
            } else if (k < 2 * HALO_W) {
                hy = HALO_W - 1;
                hx = k - HALO_W;
            } else {
                const int m = k - 2 * HALO_W;
                hy = 1 + (m >> 1);
                hx = (m & 1) ? HALO_W - 1 : 0;
            }
            rl[j] = hy * HALO_W + hx;
            const int rx = x0 - 1 + hx, ry = y0 - 1 + hy;
            if (rx >= 0 && ry >= 0 && rx < a.W && ry < a.H) {
                const size_t gi = (size_t)ry * a.W + rx;
                rs[j] = a.st[gi];
                rw[j] = a.w[gi];
            }
        }
    }
    uint32_t open = 0;
#pragma unroll
    for (int h = 0; h < PF_LOAD_BATCHES; ++h) {
        constexpr int HB = STRIP / PF_LOAD_BATCHES;
        uint8_t sv[HB];
        float wv[HB], zv[HB];
#pragma unroll
        for (int k = 0; k < HB; ++k) {
            const int gy = y0 + strip * STRIP + h * HB + k;
            sv[k] = PF_NODATA;
            wv[k] = PF_INF;
            zv[k] = 0.0f;
            if (gx < a.W && gy < a.H) {
                const size_t gi = (size_t)gy * a.W + gx;
                sv[k] = a.st[gi];
                wv[k] = a.w[gi];
                if (zq) zv[k] = zq[gi];
            }
        }
#pragma unroll
        for (int k = 0; k < HB; ++k) {
            const int ly = strip * STRIP + h * HB + k;
            S.rel[(ly + 1) * HALO_W + lx + 1] = relay(sv[k], wv[k]);
            zr[h * HB + k] = zv[k];
            if (sv[k] == PF_OPEN) open |= 1u << (h * HB + k);
        }
    }
#pragma unroll
    for (int j = 0; j < RIT; ++j)
        if (rl[j] >= 0) S.rel[rl[j]] = relay(rs[j], rw[j]);
    if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        S.bits = 0;
        S.any = 0;
        S.steps = 0;
    }
    return open;
}

// Ends a tile: OR of border bits, the next phase's list (a tile still open
// after the first pass of this phase joins phase list `phase_out`),
// neighbour activation with the rows each neighbour must solve again in
// pass `pass` + 1.  `chg` holds the thread's changed cells that its
// neighbours see (bit k = row k of its strip).
__device__ inline void finish_tile(PfShared& S, const PfArgs& a, const PfBlock& b, uint32_t bits,
                                   uint32_t open_left, uint32_t nfixed, uint32_t chg,
                                   uint32_t* cert_ctr, int* work_out, uint32_t* count_out,
                                   uint32_t* count_back, uint32_t epoch, int pass, int phase_out) {
    // per-lane: bits < 256, nfixed <= STRIP = 16
    bits = wave_or8(bits);
    open_left = __any(open_left != 0u) ? 1u : 0u;
    nfixed = wave_sum5(nfixed);
    edge_columns_put(S.wcol, S.ecol, chg);
    if ((threadIdx.x & 63) == 0) {
        if (bits) atomicOr(&S.bits, bits);
        if (open_left) atomicOr(&S.any, 1u);
        if (nfixed) atomicAdd(cert_ctr, nfixed);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.topen[b.gbase + b.lb] = S.any;
        a.tcost[b.gbase + b.lb] = S.steps;
    }
    if (pass == 0 && threadIdx.x == 0 && S.any) {
        int* lst = a.work + (size_t)(3 + phase_out) * a.total;
        if (S.steps >= PF_HEAVY)
            lst[atomicAdd(&a.ctrl[CTL_PHASE + phase_out], 1u)] = b.gbase + b.lb;
        else
            lst[a.total - 1 - (int)atomicAdd(&a.ctrl[CTL_PHASEB + phase_out], 1u)] = b.gbase + b.lb;
    }
    if (threadIdx.x < 64 && S.bits) {  // uniform within the first wave
        unsigned long long w64, e64;
        edge_columns_get(S.wcol, S.ecol, w64, e64);
        const Ranked rank{a.tcost, PF_HEAVY, count_back, a.total};
        // a neighbour without OPEN cells (its open flag 0, for good: OPEN
        // cells never reappear) cannot change and is not listed
        activate_tiles(S.bits, b.lb % b.R.ntx, b.lb / b.R.ntx, b.R.ntx, b.R.nty, b.gbase,
                       work_out, count_out, a.mark, epoch, a.rows, (pass + 1) & 1, w64, e64,
                       &rank, a.topen, &S.pfront, &S.pback);
    }
    __syncthreads();
}

// OPEN cells of a raster whose last round fixed nothing are never reached
// and keep their quantized elevation.
__device__ inline void keep_quantized(const PfRaster& R, int lb) {
    const int x0 = (lb % R.ntx) * TILE, y0 = (lb / R.ntx) * TILE;
    const int gx = x0 + (threadIdx.x & 63);
    if (gx >= R.W) return;
    const int gy0 = y0 + (threadIdx.x >> 6) * STRIP;
    uint8_t sv[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k) sv[k] = gy0 + k < R.H ? R.st[(size_t)(gy0 + k) * R.W + gx] : 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if (gy0 + k >= R.H || sv[k] != PF_OPEN) continue;
        const size_t gi = (size_t)(gy0 + k) * R.W + gx;
        R.w[gi] = R.zq[gi];
    }
}

// Fill sweeps of a tile whose relays (S.rel) are in LDS, with Z of row k of
// the thread's strip (tile row ly) from z_at(k, ly), over
// the rows of `dirty` first (a 64-bit row mask) and then every row a change
// reaches.  open / zlm: the thread's OPEN cells and those whose Z is live
// (bit k = row k of its strip); lowered / fixed / fixed_live collect its
// cells whose value dropped, that were fixed, that were fixed live; nrow
// counts the wave's row steps.
template <class ZAt>
__device__ inline void fill_sweeps(PfShared& S, unsigned long long dirty, float mn,
                                   uint32_t& open, uint32_t zlm, uint32_t& lowered,
                                   uint32_t& fixed, uint32_t& fixed_live, uint32_t& nrow,
                                   const ZAt& z_at) {
    // Row step (row ly, all 64 columns at once):
    //  1. Z-rule: an OPEN cell with a live FIXED neighbour no higher than Z
    //     is fixed at Z and, if Z is live, relays Z; along the row this
    //     spreads from a cell so fixed to an OPEN neighbour whose Z is not
    //     lower (segmented propagation over the lane masks);
    //  2. every other OPEN cell is the map v -> max(Z, min(c, v)), c = min(its
    //     value, the relays above and below); FIXED cells relay a constant
    //     (their value if live, nothing if dead).  The fixed point of the row,
    //     W(x) = max(Z, min(c, W(x-1), W(x+1))), is min(L(x), R(x)) with L, R
    //     the map chains from the left and right halo cells.
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int hl = 0, hr = HALO_W - 1;  // halo columns
    uint32_t rd = strip_rows(dirty, strip);
    ClampCarry cc;
    clamp_carry_init(cc);
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        const bool upward = sweep_upward(S.rowdirty, rd, strip, sweep);
        for (int kk = 0; kk < STRIP; ++kk) {
            const int k = upward ? (STRIP - 1 - kk) : kk;
            const int ly = strip * STRIP + k;
            if (!row_due(S.rowdirty, rd, k, ly)) continue;
            const uint32_t bit = 1u << k;
            const bool op = (open & bit) != 0;
            if (__ballot(op) == 0) continue;  // no OPEN cell: nothing in the row can change
            ++nrow;
            const int li = (ly + 1) * HALO_W + lx + 1;
            // The step's reads at once: the cell, its neighbours in the rows
            // above and below (relay and live value), the live values of its
            // row neighbours and Z.
            const float2 c = S.rel[li];
            const float2 a0 = S.rel[li - HALO_W - 1], a1 = S.rel[li - HALO_W], a2 = S.rel[li - HALO_W + 1];
            const float2 b0 = S.rel[li + HALO_W - 1], b1 = S.rel[li + HALO_W], b2 = S.rel[li + HALO_W + 1];
            const float yl = S.rel[li - 1].y, yr = S.rel[li + 1].y;
            // the row's two halo cells (one address for the whole wave)
            const float hlv = __int_as_float(__builtin_amdgcn_readfirstlane(
                __float_as_int(S.rel[(ly + 1) * HALO_W + hl].x)));
            const float hrv = __int_as_float(__builtin_amdgcn_readfirstlane(
                __float_as_int(S.rel[(ly + 1) * HALO_W + hr].x)));
            const float z = z_at(k, ly);
            const float wcur = c.x;
            const float m8y = fminf(fminf(fminf(a0.y, a1.y), fminf(a2.y, yl)),
                                    fminf(fminf(yr, b0.y), fminf(b1.y, b2.y)));
            const float m6x = fminf(fminf(fminf(a0.x, a1.x), fminf(a2.x, b0.x)), fminf(b1.x, b2.x));
            const bool zlive = op && ((zlm >> k) & 1u);
            const bool seed = op && m8y <= z;
            const float zl = lane_prev(z), zrt = lane_next(z);
            const unsigned long long sd = __ballot(seed);
            unsigned long long f = sd;
            if (sd) {
                // a step x-1 -> x needs x-1 live (lane masks shifted by one
                // lane; the end lanes get 0 from the shift)
                const unsigned long long lv = __ballot(zlive);
                f = spread_row(sd, __ballot(op && zl <= z) & (lv << 1),
                               __ballot(op && zrt <= z) & (lv >> 1));
            }
            bool ch = false;
            float lo, hi;
            if ((f >> lx) & 1ull) {
                open &= ~bit;
                fixed |= bit;
                if (wcur != z) lowered |= bit;
                if (zlive) {
                    fixed_live |= bit;
                    S.rel[li] = make_float2(z, z);
                    lo = hi = z;
                } else {
                    S.rel[li] = make_float2(PF_INF, PF_INF);
                    lo = hi = PF_INF;
                }
                ch = true;
            } else if (op) {
                lo = z;
                hi = fmaxf(z, fminf(wcur, m6x));
            } else {
                lo = hi = c.y;  // live FIXED: its value; otherwise no relay
            }
            float llo = lo, lhi = hi, rlo = lo, rhi = hi;
            scan_clamp_up(llo, lhi, cc);
            scan_clamp_down(rlo, rhi, cc);
            if (open & bit) {
                const float L = clamp3(hlv, llo, lhi);
                const float R = clamp3(hrv, rlo, rhi);
                const float nw = fminf(L, R);
                if (nw < wcur) {
                    S.rel[li].x = nw;
                    lowered |= bit;
                    ch = true;
                }
            }
            if (__ballot(ch)) {
                changed = 1;
                row_mark(S.rowdirty, rd, k, ly);
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

// Fill relaxation of one tile (with immediate fixing at Z).  Cell state
// lives in LDS (relays in rel, Z in zt); each thread keeps bit masks of its
// strip's cells only, so the row loop can run in either direction without
// per-thread arrays.
// The first pass of a round is its restart: OPEN cells start from +inf and
// all of the tile's are rewritten (in round 0 they already hold +inf).  In
// it, a raster whose last round fixed nothing is finished instead.
__device__ void fill_tile(PfShared& S, const PfArgs& args, const PfBlock& b, float mn,
                          uint32_t* cert_ctr, int* work_out, uint32_t* count_out,
                          uint32_t* count_back, uint32_t epoch, int pass, int round,
                          int phase_out) {
    const PfRaster& a = b.R;
    const unsigned long long dirty = tile_dirty_rows(args.rows, b.gbase + b.lb, pass);
    if (dirty == 0) return;  // uniform
    const bool restart = pass == 0;
    if (restart && round > 0 &&
        __builtin_amdgcn_readfirstlane(ld(&args.cert[3 * b.r + (round + 2) % 3])) == 0u) {
        keep_quantized(a, b.lb);
        return;
    }
    const int x0 = (b.lb % a.ntx) * TILE, y0 = (b.lb / a.ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gx = x0 + lx;
    ZRow zr;  // Z of the thread's cells (row k of its strip)
    uint32_t open = load_tile_own(S, a, x0, y0, restart, a.zq, zr);  // OPEN cells
    __syncthreads();
    uint32_t zlm = 0;  // OPEN cells whose Z is live
#pragma unroll
    for (int k = 0; k < STRIP; ++k)
        if (((open >> k) & 1u) && pf_live(zr[k], mn)) zlm |= 1u << k;
    clear_dirty_rows(args.rows, b.gbase + b.lb, pass);
    const uint32_t rewrite = restart && round > 0 ? open : 0u;  // stale values of the last round

    uint32_t lowered = 0, fixed = 0, fixed_live = 0;
    uint32_t nrow = 0;  // rows solved by this wave
    fill_sweeps(S, dirty, mn, open, zlm, lowered, fixed, fixed_live, nrow,
                [&](int k, int) { return zr[k]; });
    if ((threadIdx.x & 63) == 0 && nrow) atomicAdd(&S.steps, nrow);

    uint32_t bits = 0, nfixed = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const size_t gi = (size_t)(y0 + ly) * a.W + gx;
        if ((lowered | rewrite) & (1u << k))
            a.w[gi] = (fixed & (1u << k)) ? zr[k] : S.rel[(ly + 1) * HALO_W + lx + 1].x;
        if (fixed & (1u << k)) {
            a.st[gi] = (fixed_live & (1u << k)) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
            ++nfixed;
        }
        if ((lowered | fixed) & (1u << k)) bits |= border_bits(lx, ly);
    }
    finish_tile(S, args, b, bits, open, nfixed, lowered | fixed, cert_ctr, work_out, count_out,
                count_back, epoch, pass, phase_out);
}

// Certification sweeps of a tile whose converged fill values (S.rel.x) are
// in LDS, over the rows of `dirty` first and then every row a change
// reaches.  open / lm: the thread's OPEN cells with a finite value and those
// whose value is live; fixed / fixed_live collect the cells it fixes (live);
// nrow counts the wave's row steps.
__device__ inline void cert_sweeps(PfShared& S, unsigned long long dirty, uint32_t& open,
                                   uint32_t lm, uint32_t& fixed, uint32_t& fixed_live,
                                   uint32_t& nrow) {
    // Row step: seeds are OPEN cells with a live FIXED neighbour no higher
    // than their value; along the row, fixing spreads from a cell fixed live
    // to an OPEN neighbour whose value is not lower (segmented propagation
    // over the lane masks).
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t rd = strip_rows(dirty, strip);
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        const bool upward = sweep_upward(S.rowdirty, rd, strip, sweep);
        for (int kk = 0; kk < STRIP; ++kk) {
            const int k = upward ? (STRIP - 1 - kk) : kk;
            const int ly = strip * STRIP + k;
            if (!row_due(S.rowdirty, rd, k, ly)) continue;
            const uint32_t bit = 1u << k;
            const bool op = (open & bit) != 0;
            if (__ballot(op) == 0) continue;  // no OPEN cell: nothing in the row can change
            ++nrow;
            const int li = (ly + 1) * HALO_W + lx + 1;
            const float w = S.rel[li].x;
            const bool live = op && ((lm >> k) & 1u);
            const bool seed = op && min8y(S.rel, li) <= w;
            // from the left neighbour x-1 (open, live) into x, and from x+1
            const float wl = lane_prev(w), wrt = lane_next(w);
            const unsigned long long sd = __ballot(seed);
            unsigned long long f = sd;
            if (sd) {
                // a step x-1 -> x needs x-1 live (lane masks shifted by one
                // lane; the end lanes get 0 from the shift)
                const unsigned long long lv = __ballot(live);
                f = spread_row(sd, __ballot(op && wl <= w) & (lv << 1),
                               __ballot(op && wrt <= w) & (lv >> 1));
            }
            bool ch = false;
            if ((f >> lx) & 1ull) {
                open &= ~bit;
                fixed |= bit;
                if (live) {
                    fixed_live |= bit;
                    S.rel[li].y = w;
                    ch = true;
                }
            }
            if (__ballot(ch)) {
                changed = 1;
                row_mark(S.rowdirty, rd, k, ly);
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
}

// Certification of one tile after the fill has converged.
__device__ void cert_tile(PfShared& S, const PfArgs& args, const PfBlock& b, float mn,
                          uint32_t* cert_ctr, int* work_out, uint32_t* count_out,
                          uint32_t* count_back, uint32_t epoch, int pass, int phase_out) {
    const PfRaster& a = b.R;
    const unsigned long long dirty = tile_dirty_rows(args.rows, b.gbase + b.lb, pass);
    if (dirty == 0) return;  // uniform
    if (pass == 0 && __builtin_amdgcn_readfirstlane(ld(&args.topen[b.gbase + b.lb])) == 0u)
        return;  // closed since it joined the phase list
    const int x0 = (b.lb % a.ntx) * TILE, y0 = (b.lb / a.ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gx = x0 + lx;
    ZRow zr_unused;
    const uint32_t sopen = load_tile_own(S, a, x0, y0, false, nullptr, zr_unused);
    uint32_t open = 0, open_inf = 0;
    __syncthreads();  // the relays are in LDS
    clear_dirty_rows(args.rows, b.gbase + b.lb, pass);
    uint32_t lm = 0;  // OPEN cells whose (converged, fixed here) value is live
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if ((sopen >> k) & 1u) {
            const float wk = S.rel[(strip * STRIP + k + 1) * HALO_W + lx + 1].x;
            if (wk < PF_INF) {
                open |= 1u << k;
                if (pf_live(wk, mn)) lm |= 1u << k;
            } else {
                open_inf |= 1u << k;
            }
        }
    }

    uint32_t fixed = 0, fixed_live = 0;
    uint32_t nrow = 0;  // rows solved by this wave
    cert_sweeps(S, dirty, open, lm, fixed, fixed_live, nrow);
    if ((threadIdx.x & 63) == 0 && nrow) atomicAdd(&S.steps, nrow);

    uint32_t bits = 0, nfixed = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if (fixed & (1u << k)) {
            const int ly = strip * STRIP + k;
            a.st[(size_t)(y0 + ly) * a.W + gx] =
                (fixed_live & (1u << k)) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
            ++nfixed;
            if (fixed_live & (1u << k)) bits |= border_bits(lx, ly);
        }
    }
    finish_tile(S, args, b, bits, open | open_inf, nfixed, fixed_live, cert_ctr, work_out,
                count_out, count_back, epoch, pass, phase_out);
}

// Runs the passes of phase `phase` (kind 0 fill, 1 certify) until a pass
// lists nothing; returns the index of the (empty) rotating list it stopped
// at, -1 when the pass cap is hit, or PF_ABANDONED when a barrier gave up.
// The first pass reads phase list phase % 3 (and plays the part of rotating
// list `li`, the empty one the last phase stopped at); its still-open tiles
// form phase list
// (phase + 1) % 3; list (phase + 2) % 3, last read before the barrier that
// ended the phase before, is cleared for the phase after.  Fixed cells
// count into counter `slot` of their raster.
// With `compacted`, the first pass reads the compacted fill list instead.
// After a pass, the next list's front and back counts come with the
// barrier (the sums of what each workgroup listed) where the blocks are
// fewer than BAR_SUM_LIMIT; else they are read from the control words.
template <int KIND, bool MULTI>
__device__ int run_passes(PfShared& S, const PfArgs& a, int li, int phase, int slot, int round,
                          bool compacted, uint32_t& epoch, float mn1, bool ok1) {
    const int pin = phase % 3, pout = (phase + 1) % 3;
    const bool sums = (uint64_t)a.total < BAR_SUM_LIMIT;
    uint32_t next_front = 0, next_back = 0;
    for (int pass = 0; pass < a.max_passes; ++pass) {
        const int cout = (li + 1) % 3, cclr = (li + 2) % 3;
        const int lin = pass > 0 ? li : compacted ? 6 : 3 + pin;
        const int cf = pass > 0 ? CTL_COUNT + li : compacted ? CTL_COMPACT : CTL_PHASE + pin;
        const int cb = pass > 0 ? CTL_COUNTB + li : compacted ? CTL_COMPACTB : CTL_PHASEB + pin;
        const bool from_bar = pass > 0 && sums;
        const uint32_t nfront = from_bar ? next_front : ld(&a.ctrl[cf]);
        const uint32_t count = nfront + (from_bar ? next_back : ld(&a.ctrl[cb]));
        if (count == 0) return li;
        ++epoch;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.ctrl[CTL_COUNT + cclr] = 0;
            a.ctrl[CTL_COUNTB + cclr] = 0;
            a.ctrl[CTL_CLAIM + cout] = 0;  // the next pass's claims (this one's were cleared before)
            if (pass == 0) {
                a.ctrl[CTL_PHASE + (phase + 2) % 3] = 0;
                a.ctrl[CTL_PHASEB + (phase + 2) % 3] = 0;
            }
            // a certification follows the fill that last read the compacted list
            if (pass == 0 && KIND == 1) {
                a.ctrl[CTL_COMPACT] = 0;
                a.ctrl[CTL_COMPACTB] = 0;
            }
            atomicAdd(&a.ctrl[CTL_PASSES], 1u);
        }
        const int* win = a.work + (size_t)lin * a.total;
        int* wout = a.work + (size_t)cout * a.total;
        if (threadIdx.x == 0) {
            S.pfront = 0u;
            S.pback = 0u;
        }
        for (uint32_t i = blockIdx.x; i < count;) {
            uint32_t next = count;
            if (threadIdx.x == 0) next = claim_next(&a.ctrl[CTL_CLAIM + li], count);
            const PfBlock b = pf_block<MULTI>(a, ranked_entry(win, a.total, nfront, i), mn1, ok1);
            if (b.ok) {  // uniform; blocks of rasters outside the domain never open
                uint32_t* cert_ctr = &a.cert[3 * b.r + slot];
                if (KIND == 0)
                    fill_tile(S, a, b, b.mn, cert_ctr, wout, &a.ctrl[CTL_COUNT + cout],
                              &a.ctrl[CTL_COUNTB + cout], epoch, pass, round, pout);
                else
                    cert_tile(S, a, b, b.mn, cert_ctr, wout, &a.ctrl[CTL_COUNT + cout],
                              &a.ctrl[CTL_COUNTB + cout], epoch, pass, pout);
            }
            if (threadIdx.x == 0) S.next = next;
            __syncthreads();
            i = S.next;
            __syncthreads();
        }
        if (!grid_sync_sum(a.gb, 0, sums ? S.pfront : 0u, sums ? S.pback : 0u, next_front,
                           next_back))
            return PF_ABANDONED;
        li = cout;
    }
    return -1;
}

// The tiles of phase list `from` still open, into the compacted fill list
// (cleared by the certification's first pass; the caller ends with a grid
// barrier).
__device__ inline void compact_open(const PfArgs& a, int from) {
    const uint32_t nf = ld(&a.ctrl[CTL_PHASE + from]);
    const uint32_t n = nf + ld(&a.ctrl[CTL_PHASEB + from]);
    const int* src = a.work + (size_t)(3 + from) * a.total;
    int* dst = a.work + (size_t)6 * a.total;
    const uint32_t nw = (n + 63u) & ~63u;  // whole waves (the appends are wave-aggregated)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
        int g = 0;
        bool keep = false, heavy = false;
        if (i < n) {
            g = ranked_entry(src, a.total, nf, i);
            keep = a.topen[g] != 0u;
            heavy = a.tcost[g] >= PF_HEAVY;
        }
        wave_append(keep && heavy, g, dst, &a.ctrl[CTL_COMPACT]);
        wave_append_back(keep && !heavy, g, dst, &a.ctrl[CTL_COMPACTB], a.total);
    }
}

// ---- resident rounds ---------------------------------------------------------
// After round 0 only a few tiles still hold OPEN cells (some 450 of the 4096
// of a 4096^2 raster), but the rounds after it take dozens of barrier-
// separated passes that each solve a handful of tiles, so their time is the
// latency of a tile's list traffic, loads and writes.  When every such tile
// can have a workgroup of its own, the launch goes on with those workgroups
// only (the others end): each keeps its tile in LDS for all the remaining
// rounds and per pass reloads only the 260 halo cells around it, solving the
// rows next to halo cells that changed.  A pass ends at a barrier over these
// workgroups, and a phase ends after a pass in which no tile changed a
// border cell (then no halo changes and every tile is at its fixed point):
// the barrier sums the workgroups' "changed a border cell" bits.  The rules
// of each phase are those of the classic passes: the same sweeps,
// write-back and counters.

// Reloads the halo ring (the restart rule as in load_tile); returns the
// 64-bit mask of tile rows next to ring cells whose relay or live value
// changed (uniform).
__device__ inline unsigned long long load_ring(PfShared& S, const PfRaster& a, int x0, int y0,
                                               bool restart) {
    constexpr int RING = 4 * HALO_W - 4;  // 260 cells
    constexpr int IT = (RING + TILE_THREADS - 1) / TILE_THREADS;
    if (threadIdx.x < 2) S.ringrows[threadIdx.x] = 0u;
    uint8_t sv[IT];
    float wv[IT];
    int li[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        const int k = threadIdx.x + j * TILE_THREADS;
        sv[j] = PF_NODATA;
        wv[j] = PF_INF;
        li[j] = -1;
        if (k < RING) {
            int hx, hy;
            if (k < HALO_W) {
                hy = 0;
                hx = k;
            } else if (k < 2 * HALO_W) {
                hy = HALO_W - 1;
                hx = k - HALO_W;
            } else {
                const int m = k - 2 * HALO_W;
                hy = 1 + (m >> 1);
                hx = (m & 1) ? HALO_W - 1 : 0;
            }
            li[j] = hy * HALO_W + hx;
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
            if (gx >= 0 && gy >= 0 && gx < a.W && gy < a.H) {
                const size_t gi = (size_t)gy * a.W + gx;
                sv[j] = a.st[gi];
                wv[j] = a.w[gi];
            }
        }
    }
    __syncthreads();  // ringrows cleared
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < IT; ++j) {
        if (li[j] < 0) continue;
        const uint8_t st = sv[j];
        const float r = (st == PF_FIXED_LIVE || (st == PF_OPEN && !restart)) ? wv[j] : PF_INF;
        const float2 nv = make_float2(r, st == PF_FIXED_LIVE ? r : PF_INF);
        const float2 ov = S.rel[li[j]];
        if (nv.x != ov.x || nv.y != ov.y) {
            S.rel[li[j]] = nv;
            const int hy = li[j] / HALO_W;
            int r0, r1;  // tile rows whose cells see this halo cell
            if (hy == 0) {
                r0 = r1 = 0;
            } else if (hy == HALO_W - 1) {
                r0 = r1 = TILE - 1;
            } else {
                r0 = max(hy - 2, 0);
                r1 = min(hy, TILE - 1);
            }
            for (int r2 = r0; r2 <= r1; ++r2) {
                if (r2 < 32) lo |= 1u << r2;
                else hi |= 1u << (r2 - 32);
            }
        }
    }
    if (lo) atomicOr(&S.ringrows[0], lo);
    if (hi) atomicOr(&S.ringrows[1], hi);
    __syncthreads();
    return ((unsigned long long)S.ringrows[1] << 32) | S.ringrows[0];
}

// Sum over the workgroup of a per-thread count below 32 and OR of 8-bit
// border masks, through LDS words (S.steps / S.bits); uniform results.
__device__ inline void resident_reduce(PfShared& S, uint32_t nfixed, uint32_t bits,
                                       uint32_t* cert_ctr, uint32_t& any_border) {
    if (threadIdx.x == 0) {
        S.bits = 0;
        S.steps = 0;
    }
    __syncthreads();
    bits = wave_or8(bits);
    nfixed = wave_sum5(nfixed);
    if ((threadIdx.x & 63) == 0) {
        if (bits) atomicOr(&S.bits, bits);
        if (nfixed) atomicAdd(&S.steps, nfixed);
    }
    __syncthreads();
    if (threadIdx.x == 0 && S.steps) atomicAdd(cert_ctr, S.steps);
    any_border = S.bits;
    __syncthreads();
}

// The rounds after round `round` for the tile of workgroup `blockIdx.x`
// (one raster; workgroups 0 .. n-1 take the n tiles of the compacted list).
// Returns the round it ended in, -1 when a phase hit the pass cap, or
// PF_ABANDONED.
__device__ int resident_rounds(PfSharedRes& S, const PfArgs& a, float mn, int round, uint32_t n) {
    const PfRaster& R = a.one;
    const uint32_t nf = ld(&a.ctrl[CTL_COMPACT]);
    const int t = ranked_entry(a.work + (size_t)6 * a.total, a.total, nf, blockIdx.x);
    const int x0 = (t % R.ntx) * TILE, y0 = (t / R.ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gx = x0 + lx;
    // the whole tile once, as a first pass of a round loads it
    load_tile(S, R, x0, y0, true);
    uint32_t open = 0, zlm = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int gy = y0 + strip * STRIP + k;
        uint8_t sv = PF_NODATA;
        float zv = 0.0f;
        if (gx < R.W && gy < R.H) {
            const size_t gi = (size_t)gy * R.W + gx;
            sv = R.st[gi];
            zv = R.zq[gi];
        }
        if (sv == PF_OPEN) {
            open |= 1u << k;
            if (pf_live(zv, mn)) zlm |= 1u << k;
        }
        S.zt[(strip * STRIP + k) * TILE + lx] = zv;
    }
    __syncthreads();
    uint32_t q = 0;  // resident pass count
    bool first_round = true;
    for (++round;; ++round) {
        const int slot = round % 3;
        // a round after one that fixed nothing: the OPEN cells are never
        // reached and keep Z
        if (__builtin_amdgcn_readfirstlane(ld(&a.cert[(round + 2) % 3])) == 0u) {
#pragma unroll
            for (int k = 0; k < STRIP; ++k) {
                const int gy = y0 + strip * STRIP + k;
                if ((open >> k) & 1u) R.w[(size_t)gy * R.W + gx] = S.zt[(strip * STRIP + k) * TILE + lx];
            }
            return round;
        }
        // ---- fill: OPEN cells restart from +inf and are all rewritten
        if (!first_round) {
#pragma unroll
            for (int k = 0; k < STRIP; ++k)
                if ((open >> k) & 1u)
                    S.rel[(strip * STRIP + k + 1) * HALO_W + lx + 1] = make_float2(PF_INF, PF_INF);
        }
        first_round = false;
        uint32_t rewrite = open;
        for (int pass = 0;; ++pass, ++q) {
            if (pass >= a.max_passes) return -1;
            unsigned long long dirty = ~0ull;
            const unsigned long long ring = load_ring(S, R, x0, y0, pass == 0);
            if (pass > 0) dirty = ring;
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.ctrl[CTL_PASSES], 1u);
            uint32_t any_border = 0;
            if (dirty) {
                if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
                __syncthreads();
                uint32_t lowered = 0, fixed = 0, fixed_live = 0, nrow = 0;
                // the carried scans' 20 registers push 3 values of the rounds'
                // last write-back to scratch here, off the sweep loop
                fill_sweeps(S, dirty, mn, open, zlm, lowered, fixed, fixed_live, nrow,
                            [&](int, int ly) { return S.zt[ly * TILE + lx]; });
                uint32_t bits = 0, nfixed = 0;
#pragma unroll
                for (int k = 0; k < STRIP; ++k) {
                    const int ly = strip * STRIP + k;
                    const size_t gi = (size_t)(y0 + ly) * R.W + gx;
                    if ((lowered | rewrite) & (1u << k))
                        R.w[gi] = (fixed & (1u << k)) ? S.zt[ly * TILE + lx]
                                                      : S.rel[(ly + 1) * HALO_W + lx + 1].x;
                    if (fixed & (1u << k)) {
                        R.st[gi] = (fixed_live & (1u << k)) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
                        ++nfixed;
                    }
                    if ((lowered | fixed) & (1u << k)) bits |= border_bits(lx, ly);
                }
                rewrite = 0;
                resident_reduce(S, nfixed, bits, &a.cert[slot], any_border);
            }
            uint32_t changed = 0, unused = 0;
            if (!grid_sync_sum(a.gb, n, any_border ? 1u : 0u, 0u, changed, unused))
                return PF_ABANDONED;
            if (!changed) {
                ++q;
                break;
            }
        }
        // the next round's fixed-cell counter: last read at this round's start
        if (blockIdx.x == 0 && threadIdx.x == 0) a.cert[(round + 1) % 3] = 0;
        // ---- certification of the converged fill
        uint32_t copen = 0, lm = 0, open_inf = 0;
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            if (!((open >> k) & 1u)) continue;
            const float wk = S.rel[(strip * STRIP + k + 1) * HALO_W + lx + 1].x;
            if (wk < PF_INF) {
                copen |= 1u << k;
                if (pf_live(wk, mn)) lm |= 1u << k;
            } else {
                open_inf |= 1u << k;
            }
        }
        for (int pass = 0;; ++pass, ++q) {
            if (pass >= a.max_passes) return -1;
            const unsigned long long ring = load_ring(S, R, x0, y0, false);
            const unsigned long long dirty = pass == 0 ? ~0ull : ring;
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a.ctrl[CTL_PASSES], 1u);
            uint32_t any_border = 0;
            if (dirty) {
                if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
                __syncthreads();
                uint32_t fixed = 0, fixed_live = 0, nrow = 0;
                cert_sweeps(S, dirty, copen, lm, fixed, fixed_live, nrow);
                uint32_t bits = 0, nfixed = 0;
#pragma unroll
                for (int k = 0; k < STRIP; ++k) {
                    if (!(fixed & (1u << k))) continue;
                    const int ly = strip * STRIP + k;
                    const bool live = (fixed_live >> k) & 1u;
                    R.st[(size_t)(y0 + ly) * R.W + gx] = live ? PF_FIXED_LIVE : PF_FIXED_DEAD;
                    // a cell fixed dead relays nothing from now on
                    if (!live)
                        S.rel[(ly + 1) * HALO_W + lx + 1] = make_float2(PF_INF, PF_INF);
                    ++nfixed;
                    if (live) bits |= border_bits(lx, ly);
                }
                open &= ~fixed;
                resident_reduce(S, nfixed, bits, &a.cert[slot], any_border);
            }
            uint32_t changed = 0, unused = 0;
            if (!grid_sync_sum(a.gb, n, any_border ? 1u : 0u, 0u, changed, unused))
                return PF_ABANDONED;
            if (!changed) {
                ++q;
                break;
            }
        }
        (void)open_inf;
    }
}

// MULTI: a table of rasters (else the one raster of the arguments; two
// instantiations keep the one-raster code free of table lookups).  At least
// three waves a SIMD; the one-raster kernel (126 VGPRs, 35.4 KB of LDS) runs
// four workgroups a CU, the table one (130 VGPRs) three (the one-raster
// kernel times the same at three and at four).
template <bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_pf_persistent(PfArgs a) {
    __shared__ PfShared S;
    if (grid_aborted(a.gb)) return;

    // One raster: its minimum and domain flag, once.
    float mn1 = 0.0f;
    bool ok1 = true;
    if (!MULTI) ok1 = pf_raster_ok(a.one, mn1);

    // Initial states (PitFillingAlgorithmCPUOpenMP.h:121-144 seeds) in units
    // of UNIT_ROWS x 64 cells, one wave each.  A raster outside the domain
    // gets its status and no work: its blocks are skipped and never open.
    {
        const int lane = threadIdx.x & 63;
        const long long units = (long long)a.total * UNITS_PER_BLOCK;
        for (long long u = wave_unit_first(); u < units; u += wave_unit_stride()) {
            const int g = (int)(u / UNITS_PER_BLOCK);
            const int row0 = (int)(u % UNITS_PER_BLOCK) * UNIT_ROWS;
            const PfBlock b = pf_block<MULTI>(a, g, mn1, ok1);
            const PfRaster& R = b.R;
            float mn = mn1;
            bool ok = ok1;
            if (MULTI) {
                ok = pf_raster_ok(R, mn);  // the per-raster words are written here
                if (b.lb == 0 && row0 == 0 && lane == 0)
                    a.rinfo[b.r] = make_uint2(__float_as_uint(mn), ok ? 1u : 0u);
            }
            if (!ok) {
                if (b.lb == 0 && row0 == 0 && lane == 0) *R.status = (uint32_t)(-CAGEO_ERR_RANGE);
                continue;
            }
            const int x = (b.lb % R.ntx) * TILE + lane, y0 = (b.lb / R.ntx) * TILE + row0;
            float z[UNIT_ROWS];
#pragma unroll
            for (int k = 0; k < UNIT_ROWS; ++k)
                z[k] = (x < R.W && y0 + k < R.H) ? R.zq[(size_t)(y0 + k) * R.W + x] : 0.0f;
#pragma unroll
            for (int k = 0; k < UNIT_ROWS; ++k) {
                const int y = y0 + k;
                if (x >= R.W || y >= R.H) continue;
                uint8_t s;
                float v;
                if (z[k] == R.nodata) {
                    s = PF_NODATA;
                    v = z[k];
                } else if (x == 0 || y == 0 || x == R.W - 1 || y == R.H - 1 || z[k] == 0.0f) {
                    s = pf_live(z[k], mn) ? PF_FIXED_LIVE : PF_FIXED_DEAD;
                    v = z[k];
                } else {
                    s = PF_OPEN;
                    v = PF_INF;
                }
                const size_t i = (size_t)y * R.W + x;
                R.st[i] = s;
                R.w[i] = v;
            }
        }
    }
    // phase list 0: every block (those of rasters outside the domain are skipped)
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < a.total; g += gridDim.x * blockDim.x) {
        a.work[(size_t)3 * a.total + g] = g;
        a.topen[g] = 1u;  // until its first solve
        a.tcost[g] = 0u;
        reinterpret_cast<uint4*>(a.rows)[g] = make_uint4(0u, 0u, 0u, 0u);
        // an abandoned launch may have stamped this block with an epoch the
        // next launch reuses (the counter advances when a launch completes)
        a.mark[g] = 0u;
    }
    for (int k = threadIdx.x; k < 3 * a.nr; k += blockDim.x)
        if (blockIdx.x == 0) a.cert[k] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl[CTL_COUNT + 0] = 0;
        a.ctrl[CTL_COUNT + 1] = 0;
        a.ctrl[CTL_COUNT + 2] = 0;
        a.ctrl[CTL_PHASE + 0] = (uint32_t)a.total;
        a.ctrl[CTL_PHASE + 1] = 0;
        a.ctrl[CTL_PHASE + 2] = 0;
        a.ctrl[CTL_CLAIM + 0] = 0;
        a.ctrl[CTL_CLAIM + 1] = 0;
        a.ctrl[CTL_CLAIM + 2] = 0;
        a.ctrl[CTL_COMPACT] = 0;
        a.ctrl[CTL_COUNTB + 0] = 0;
        a.ctrl[CTL_COUNTB + 1] = 0;
        a.ctrl[CTL_COUNTB + 2] = 0;
        a.ctrl[CTL_PHASEB + 0] = 0;
        a.ctrl[CTL_PHASEB + 1] = 0;
        a.ctrl[CTL_PHASEB + 2] = 0;
        a.ctrl[CTL_COMPACTB] = 0;
        a.ctrl[CTL_RESIDENT] = 0;
        a.ctrl[CTL_PASSES] = 0;  // written by workgroup 0 only, read by the host
    }
    uint32_t epoch = ld(&a.ctrl[CTL_EPOCH]);
    if (!grid_sync(a.gb)) return;

    // Phases alternate fill and certification.  Round r counts fixed cells
    // into counter slot r % 3 of each raster; its fill's first pass reads
    // slot (r+2) % 3 (round r-1's); slot (r+1) % 3, last read by round r-1's
    // first pass, is cleared once round r's fill is over.
    int li = 0, phase = 0, round = 0;
    for (;; ++round) {
        const int slot = round % 3;
        li = run_passes<0, MULTI>(S, a, li, phase, slot, round, round > 0, epoch, mn1, ok1);
        if (li == PF_ABANDONED) return;
        if (li < 0) break;
        ++phase;
        if (blockIdx.x == 0)
            for (int r = threadIdx.x; r < a.nr; r += blockDim.x) a.cert[3 * r + (round + 1) % 3] = 0;
        if (ld(&a.ctrl[CTL_PHASE + phase % 3]) + ld(&a.ctrl[CTL_PHASEB + phase % 3]) == 0) break;
        li = run_passes<1, MULTI>(S, a, li, phase, slot, round, false, epoch, mn1, ok1);
        if (li == PF_ABANDONED) return;
        if (li < 0) break;
        ++phase;
        if (ld(&a.ctrl[CTL_PHASE + phase % 3]) + ld(&a.ctrl[CTL_PHASEB + phase % 3]) == 0) break;
        compact_open(a, phase % 3);
        if (!grid_sync(a.gb)) return;
        const uint32_t nopen = ld(&a.ctrl[CTL_COMPACT]) + ld(&a.ctrl[CTL_COMPACTB]);
        if (nopen == 0) break;
        if (!MULTI && nopen <= (uint32_t)a.resident_cap) {
            // every tile still open can have a workgroup of its own: the
            // rounds go on in k_pf_resident, launched after this kernel
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                a.ctrl[CTL_RESIDENT] = nopen;
                a.ctrl[CTL_RROUND] = (uint32_t)round;
                a.ctrl[CTL_EPOCH] = epoch;
            }
            return;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl[CTL_EPOCH] = epoch;
        a.ctrl[CTL_ROUNDS] = (uint32_t)round + 1;
        if (li < 0) a.ctrl[CTL_STATUS] = (uint32_t)(-CAGEO_ERR_NOT_CONVERGED);
    }
}

// The rounds after the one k_pf_persistent stopped at, when it handed them
// over (ctrl[CTL_RESIDENT] open tiles, one workgroup each; the other
// workgroups end at once).  Nothing to do otherwise.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pf_resident(PfArgs a) {
    __shared__ PfSharedRes S;
    const uint32_t n = __builtin_amdgcn_readfirstlane(ld(&a.ctrl[CTL_RESIDENT]));
    if (blockIdx.x >= n) return;
    if (grid_aborted(a.gb)) return;
    float mn = 0.0f;
    pf_raster_ok(a.one, mn);  // in the domain: the first kernel checked it
    const int round = (int)__builtin_amdgcn_readfirstlane(ld(&a.ctrl[CTL_RROUND]));
    const int rr = resident_rounds(S, a, mn, round, n);
    if (rr == PF_ABANDONED) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl[CTL_ROUNDS] = (uint32_t)(rr < 0 ? round : rr) + 1;
        if (rr < 0) a.ctrl[CTL_STATUS] = (uint32_t)(-CAGEO_ERR_NOT_CONVERGED);
    }
}

PfRaster pf_raster(Context* c, const float* zq, float* w, float nodata) {
    PfRaster R;
    R.zq = zq;
    R.w = w;
    R.st = c->d_st;
    R.W = c->W;
    R.H = c->H;
    R.ntx = c->ntx;
    R.nty = c->nty;
    R.nodata = nodata;
    R.scal = c->d_scal;
    R.status = c->d_ctrl + CTL_STATUS;
    return R;
}

int occupancy_cap(const void* fn, int device, int* cap) {
    int per_cu = 0;
    CAGEO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TILE_THREADS, 0));
    hipDeviceProp_t prop;
    CAGEO_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (per_cu < 1) return CAGEO_ERR_HIP;
    *cap = per_cu * prop.multiProcessorCount;
    return CAGEO_OK;
}

}  // namespace

// Co-resident workgroups of the persistent kernel on this device: the most
// its grid barrier allows.
int pit_fill_grid_cap(Context* c, int* cap) {
    return occupancy_cap(reinterpret_cast<const void*>(k_pf_persistent<false>), c->device, cap);
}

int pit_fill_resident_cap(Context* c, int* cap) {
    return occupancy_cap(reinterpret_cast<const void*>(k_pf_resident), c->device, cap);
}

int pit_fill_persistent_grid(Context* c, int* grid) {
    if (c->pf_grid <= 0) CAGEO_TRY(pit_fill_grid_cap(c, &c->pf_grid));
    *grid = c->pf_grid;
    return CAGEO_OK;
}

int pit_fill_quantized(Context* c, float* zq, float* w, float nodata, bool /*stats_ready*/) {
    int grid = 0;
    CAGEO_TRY(pit_fill_persistent_grid(c, &grid));
    PfArgs a;
    a.one = pf_raster(c, zq, w, nodata);
    a.rs = nullptr;
    a.bmap = nullptr;
    a.rinfo = nullptr;
    a.nr = 1;
    a.total = c->ntiles;
    a.work = c->d_work3;
    a.mark = c->d_mark_pf;
    a.topen = c->d_topen;
    a.tcost = c->d_tcost;
    a.rows = c->d_rows;
    a.ctrl = c->d_ctrl;
    a.cert = c->d_ctrl + CTL_CERT;
    a.gb = context_gridbar(c);
    a.max_passes = 1 << 20;
    if (c->pf_res_grid <= 0) CAGEO_TRY(pit_fill_resident_cap(c, &c->pf_res_grid));
    a.resident_cap = std::min(c->pf_res_grid, grid);
    // Ordinary launches: the grid barrier needs its workgroups co-resident,
    // which the occupancy caps guarantee (work of other kernels on the device
    // can only delay the start of some workgroups).  Cooperative launches are
    // not used: the runtime allows them to one process at a time.  The second
    // launch goes on with the rounds the first one hands over, if any.
    hipLaunchKernelGGL(k_pf_persistent<false>, dim3(grid), dim3(TILE_THREADS), 0, c->stream, a);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pf_resident, dim3(a.resident_cap), dim3(TILE_THREADS), 0, c->stream, a);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

// Raster table of a batch: each context's d_z -> d_w with its DEM nodata.
int pit_fill_tables(MultiState* m, Context* const* ctxs) {
    std::vector<PfRaster> t(m->nr);
    for (int r = 0; r < m->nr; ++r) t[r] = pf_raster(ctxs[r], ctxs[r]->d_z, ctxs[r]->d_w, ctxs[r]->nodata_dem);
    CAGEO_HIP_TRY(hipMalloc(&m->d_pf, t.size() * sizeof(PfRaster)));
    CAGEO_HIP_TRY(hipMemcpy(m->d_pf, t.data(), t.size() * sizeof(PfRaster), hipMemcpyHostToDevice));
    CAGEO_TRY(occupancy_cap(reinterpret_cast<const void*>(k_pf_persistent<true>),
                            ctxs[0]->device, &m->pf_cap));
    if (m->pf_grid <= 0) m->pf_grid = m->pf_cap;
    return CAGEO_OK;
}

int pit_fill_multi(MultiState* m, hipStream_t s) {
    PfArgs a;
    a.one = PfRaster{};
    a.rs = static_cast<const PfRaster*>(m->d_pf);
    a.bmap = m->d_bmap;
    a.rinfo = m->d_rinfo;
    a.nr = m->nr;
    a.total = m->total;
    a.work = m->d_work;
    a.mark = m->d_mark_pf;
    a.topen = m->d_topen;
    a.tcost = m->d_tcost;
    a.rows = m->d_rows;
    a.ctrl = m->d_ctrl;
    a.cert = m->d_cert;
    a.gb = multi_gridbar(m);
    a.max_passes = 1 << 20;
    a.resident_cap = 0;  // the batch keeps every round in the all-tile launch
    hipLaunchKernelGGL(k_pf_persistent<true>, dim3(m->pf_grid), dim3(TILE_THREADS), 0, s, a);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
