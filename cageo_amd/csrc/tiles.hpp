// Tile worklists for the iterative stencil stages (pit fill, certification,
// flat resolution).  In a pass, each listed 64x64 tile is taken by one
// 256-thread workgroup of the persistent kernel; the tile iterates to a local
// fixed point in LDS against the halo it loaded, writes back, and lists the
// neighbour tiles whose halo it changed (border cells only) for the next
// pass.  A per-tile epoch stamp de-duplicates the next list without clearing
// it.  The passes stop when one lists nothing: then every tile is at a fixed
// point with respect to its neighbours' final values.
#pragma once

#include "common.hpp"

namespace cageo {

// Neighbour-tile bits: N, NE, E, SE, S, SW, W, NW.
__device__ constexpr int TDX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ constexpr int TDY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};
// Index d of the neighbour tile at offset (ox, oy), TDX[d] == ox, TDY[d] == oy.
// (A table of nibbles in a register: an indexed array would live in
// private memory.)
__device__ inline int tile_dir(int ox, int oy) {
    return (int)((0x345206107ull >> (4 * ((oy + 1) * 3 + ox + 1))) & 7u);
}

// ---- exact row solves of the LDS tile sweeps -------------------------------
// A wave owns a strip of STRIP tile rows, one lane per column.  Each step
// solves one row exactly with its two neighbour rows held at their current
// values: along the row the relaxation is a chain of per-cell monotone maps,
// composed for all 64 columns at once by wave scans (6 DPP steps in each
// direction), so information crosses the whole row in one step rather than
// one column per sweep.  Rows are visited in alternating order (Gauss-
// Seidel down, then up); a row whose cells changed marks its neighbour rows
// for another solve, and a tile is done after a sweep without changes.

// min(max(v, lo), hi) for lo <= hi and no NaN: the median of the three (one
// exact selection, v_med3_f32).
__device__ inline float clamp3(float v, float lo, float hi) {
    return __builtin_amdgcn_fmed3f(v, lo, hi);
}

// Clamp map v -> min(max(v, lo), hi) with lo <= hi.  Replaces (lo, hi) by
// the composition "self after prev".  Compositions of clamps are clamps
// (and keep lo <= hi).
__device__ inline void clamp_after(float& lo, float& hi, float plo, float phi) {
    const float nlo = clamp3(plo, lo, hi);
    const float nhi = clamp3(phi, lo, hi);
    lo = nlo;
    hi = nhi;
}

// ---- lane movement on the VALU (DPP) -----------------------------------------
// GFX9 data-parallel-primitive controls: row shifts within 16-lane rows,
// whole-wave shifts by one lane, broadcasts of lane 15 / 31 into the next
// rows.  A lane whose source is outside the pattern, or whose row is masked
// off, keeps `old`.  These replace LDS permutes (ds_bpermute) in the row
// solves: no LDS traffic and a few cycles of latency per step.
constexpr int DPP_ROW_SHL = 0x100;  // + d: lane x reads x + d within its row
constexpr int DPP_ROW_SHR = 0x110;  // + d: lane x reads x - d within its row
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROWS = 0xf>
__device__ inline int dpp_i(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ inline float dpp_f(float old, float src) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, ROWS, 0xf, false));
}
// Value of lane x - 1 (lane 0 keeps its own) and of lane x + 1 (lane 63 keeps
// its own), as __shfl_up / __shfl_down by one.
__device__ inline float lane_prev(float v) { return dpp_f<DPP_WAVE_SHR1>(v, v); }
__device__ inline float lane_next(float v) { return dpp_f<DPP_WAVE_SHL1>(v, v); }
__device__ inline int lane_prev(int v) { return dpp_i<DPP_WAVE_SHR1>(v, v); }
__device__ inline int lane_next(int v) { return dpp_i<DPP_WAVE_SHL1>(v, v); }
__device__ inline float lane_read(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Inclusive scan of an associative "self after prev" composition C(self,
// prev) with identity element E, over the 64 lanes of a wave: lane x ends
// with f_x o ... o f_0 (row shifts, then the row broadcasts carry rows 0 /
// 0-1 into the rows after them).  The row solves use the carried form below,
// which also scans towards lower lanes (lane x ends with f_x o ... o f_63:
// row shifts, then the totals of the rows after each row, read from lanes
// 16, 32, 48).
#define CAGEO_SCAN_UP(DPP, a, b, ea, eb, C)                                                    \
    do {                                                                                       \
        C(a, b, DPP<DPP_ROW_SHR + 1>(ea, a), DPP<DPP_ROW_SHR + 1>(eb, b));                     \
        C(a, b, DPP<DPP_ROW_SHR + 2>(ea, a), DPP<DPP_ROW_SHR + 2>(eb, b));                     \
        C(a, b, DPP<DPP_ROW_SHR + 4>(ea, a), DPP<DPP_ROW_SHR + 4>(eb, b));                     \
        C(a, b, DPP<DPP_ROW_SHR + 8>(ea, a), DPP<DPP_ROW_SHR + 8>(eb, b));                     \
        C(a, b, (DPP<DPP_ROW_BCAST15, 0xa>(ea, a)), (DPP<DPP_ROW_BCAST15, 0xa>(eb, b)));       \
        C(a, b, (DPP<DPP_ROW_BCAST31, 0xc>(ea, a)), (DPP<DPP_ROW_BCAST31, 0xc>(eb, b)));       \
    } while (0)
__device__ inline int lane_read(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }


// Scans with their lane moves' destinations carried across row steps.
// Each destination is written by its own lane move only, so the lanes that
// move leaves alone (outside the shift, or in rows it masks off) keep the
// identity they were given once, and no step has to write it again (20
// moves a row step fewer, for 20 registers).
template <class T>
struct ScanCarry {
    T u[12];  // upward scan: levels 1, 2, 4, 8, bcast15, bcast31 (first, second)
    T d[8];   // downward scan: levels 1, 2, 4, 8 (first, second)
};
template <class T>
__device__ inline void scan_carry_init(ScanCarry<T>& c, T ea, T eb) {
#pragma unroll
    for (int i = 0; i < 12; i += 2) {
        c.u[i] = ea;
        c.u[i + 1] = eb;
    }
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
        c.d[i] = ea;
        c.d[i + 1] = eb;
    }
}
template <int CTRL, int ROWS>
__device__ inline float dpp_t(float old, float src) { return dpp_f<CTRL, ROWS>(old, src); }
template <int CTRL, int ROWS>
__device__ inline int dpp_t(int old, int src) { return dpp_i<CTRL, ROWS>(old, src); }
#define CAGEO_CARRY_STEP(ARR, I, CTRL, ROWS)                                                   \
    do {                                                                                       \
        ARR[I] = dpp_t<CTRL, ROWS>(ARR[I], a);                                                 \
        ARR[I + 1] = dpp_t<CTRL, ROWS>(ARR[I + 1], b);                                         \
        compose(a, b, ARR[I], ARR[I + 1]);                                                     \
    } while (0)
// Inclusive scans of "self after prev" (compose(self_a, self_b, prev_a,
// prev_b)) with identity (ea, eb), towards higher and towards lower lanes.
template <class T, class C>
__device__ inline void scan_up_carried(T& a, T& b, ScanCarry<T>& c, C compose) {
    CAGEO_CARRY_STEP(c.u, 0, DPP_ROW_SHR + 1, 0xf);
    CAGEO_CARRY_STEP(c.u, 2, DPP_ROW_SHR + 2, 0xf);
    CAGEO_CARRY_STEP(c.u, 4, DPP_ROW_SHR + 4, 0xf);
    CAGEO_CARRY_STEP(c.u, 6, DPP_ROW_SHR + 8, 0xf);
    CAGEO_CARRY_STEP(c.u, 8, DPP_ROW_BCAST15, 0xa);
    CAGEO_CARRY_STEP(c.u, 10, DPP_ROW_BCAST31, 0xc);
}
template <class T, class C>
__device__ inline void scan_down_carried(T& a, T& b, ScanCarry<T>& c, T ea, T eb, C compose) {
    CAGEO_CARRY_STEP(c.d, 0, DPP_ROW_SHL + 1, 0xf);
    CAGEO_CARRY_STEP(c.d, 2, DPP_ROW_SHL + 2, 0xf);
    CAGEO_CARRY_STEP(c.d, 4, DPP_ROW_SHL + 4, 0xf);
    CAGEO_CARRY_STEP(c.d, 6, DPP_ROW_SHL + 8, 0xf);
    T s3a = lane_read(a, 48), s3b = lane_read(b, 48);
    T s2a = lane_read(a, 32), s2b = lane_read(b, 32);
    T s1a = lane_read(a, 16), s1b = lane_read(b, 16);
    compose(s2a, s2b, s3a, s3b);
    compose(s1a, s1b, s2a, s2b);
    const int row = (threadIdx.x & 63) >> 4;
    compose(a, b, row == 0 ? s1a : row == 1 ? s2a : row == 2 ? s3a : ea,
            row == 0 ? s1b : row == 1 ? s2b : row == 2 ? s3b : eb);
}
#undef CAGEO_CARRY_STEP

typedef ScanCarry<float> ClampCarry;
__device__ inline void clamp_carry_init(ClampCarry& c) {
    scan_carry_init(c, -__builtin_huge_valf(), __builtin_huge_valf());
}
__device__ inline void scan_clamp_up(float& lo, float& hi, ClampCarry& c) {
    scan_up_carried(lo, hi, c, clamp_after);
}
__device__ inline void scan_clamp_down(float& lo, float& hi, ClampCarry& c) {
    scan_down_carried(lo, hi, c, -__builtin_huge_valf(), __builtin_huge_valf(), clamp_after);
}

// Rows due for a row solve.  Each wave keeps the rows of its own strip in
// one wave-uniform word (bit k = row k of the strip): the low half holds
// rows reached by a change below them (an upward sweep carries it on), the
// high half rows reached by a change above them; a row due regardless of
// direction is in both.  The first and last rows of a strip are also
// marked by the neighbouring strips' waves, through S.rowdirty.  A wave
// clears such a flag before it reads the row's neighbours, and a marking
// wave writes its row before the flag (LDS operations of a wave complete in
// order), so a change is either seen by the solve or flagged for the next
// sweep.  The sweeps take the strip index through readfirstlane: the
// compiler cannot tell that threadIdx.x >> 6 is the same across a wave, and
// would otherwise keep this word, the row index and every test on them in
// vector registers behind divergent branches (the pit fill's persistent
// kernel: 168 VGPRs and a spill before, 106 and none after; 4.8 -> 4.6 ms).
static_assert(STRIP == 16, "two strip masks share a word");
__device__ inline uint32_t strip_rows(unsigned long long dirty, int strip) {
    const uint32_t m = (uint32_t)(dirty >> (strip * STRIP)) & 0xffffu;
    return m | (m << 16);
}
// Direction of the next sweep (true: upward): the one its pending rows ask
// for, a strip edge flagged by the strip above asking for a downward sweep
// and one flagged by the strip below for an upward one; alternating when
// both or neither are pending.
__device__ inline bool sweep_upward(const uint32_t* rowdirty, uint32_t rd, int strip, int sweep) {
    const bool top = __builtin_amdgcn_readfirstlane(rowdirty[strip * STRIP]) != 0u;
    const bool bottom = __builtin_amdgcn_readfirstlane(rowdirty[strip * STRIP + STRIP - 1]) != 0u;
    const bool dn = (rd >> 16) != 0u || top, up = (rd & 0xffffu) != 0u || bottom;
    return dn != up ? up : (sweep & 1) != 0;
}
__device__ inline bool row_due(uint32_t* rowdirty, uint32_t& rd, int k, int ly) {
    const uint32_t bits = 0x10001u << k;
    bool due = (rd & bits) != 0u;
    rd &= ~bits;
    if (k == 0 || k == STRIP - 1) {
        if (__builtin_amdgcn_readfirstlane(rowdirty[ly]) != 0u) {
            rowdirty[ly] = 0u;
            due = true;
        }
    }
    return due;
}
// Row ly (row k of the strip) changed in a way that concerns the row above
// (`up`) and/or the row below (`down`): those become due again.
__device__ inline void row_mark_dirs(uint32_t* rowdirty, uint32_t& rd, int k, int ly, bool up,
                                     bool down) {
    if (up) {
        if (k > 0)
            rd |= 1u << (k - 1);
        else if (ly > 0 && (threadIdx.x & 63) == 0)
            rowdirty[ly - 1] = 1u;
    }
    if (down) {
        if (k < STRIP - 1)
            rd |= 1u << (k + 17);
        else if (ly < TILE - 1 && (threadIdx.x & 63) == 0)
            rowdirty[ly + 1] = 1u;
    }
}

// Marks the rows next to row k of the strip (row ly of the tile).
__device__ inline void row_mark(uint32_t* rowdirty, uint32_t& rd, int k, int ly) {
    if (k > 0)
        rd |= 1u << (k - 1);
    else if (ly > 0 && (threadIdx.x & 63) == 0)
        rowdirty[ly - 1] = 1u;
    if (k < STRIP - 1)
        rd |= 1u << (k + 17);
    else if (ly < TILE - 1 && (threadIdx.x & 63) == 0)
        rowdirty[ly + 1] = 1u;
}

// Tiles of a pass are taken in turn: the first gridDim.x by workgroup
// number, the rest claimed from counter `claim` (a slow tile holds back no
// tile behind it).  The claim is issued when the workgroup starts a tile
// and read when it ends, so its latency hides behind the tile's loads.
// Called by thread 0; returns the workgroup's next index (>= count: none).
__device__ inline uint32_t claim_next(uint32_t* claim, uint32_t count) {
    return count > gridDim.x ? gridDim.x + atomicAdd(claim, 1u) : count;
}

// Dirty-row masks of tiles solved in passes (two 32-bit words per pass
// parity, four per tile, at `rows`).  The rows of global tile g due for a
// solve in pass `pass` of a phase: every row in the phase's first pass;
// later, the rows its neighbours' changes unsettled in the previous pass
// (every other row is still at the fixed point of its last solve, its halo
// unchanged).  Zero: nothing to do.  Wave-uniform.
__device__ inline unsigned long long tile_dirty_rows(const uint32_t* rows, int g, int pass) {
    if (pass == 0) return ~0ull;
    const uint32_t* rw = rows + 4 * (size_t)g + 2 * (pass & 1);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(rw[0]);
    const uint32_t hi = __builtin_amdgcn_readfirstlane(rw[1]);
    return ((unsigned long long)hi << 32) | lo;
}
// Clears what tile_dirty_rows read; called once every thread has read it.
__device__ inline void clear_dirty_rows(uint32_t* rows, int g, int pass) {
    if (pass > 0 && threadIdx.x == 0) {
        uint32_t* rw = rows + 4 * (size_t)g + 2 * (pass & 1);
        rw[0] = 0u;
        rw[1] = 0u;
    }
}
// Changed rows of the tile's columns 0 and TILE-1, from each thread's mask
// `chg` of changed cells (bit k = row k of its strip): put by every thread
// before a __syncthreads, got as 64-bit row masks after it.
__device__ inline void edge_columns_put(uint32_t* wcol, uint32_t* ecol, uint32_t chg) {
    const uint32_t wc = __builtin_amdgcn_readlane(chg, 0);
    const uint32_t ec = __builtin_amdgcn_readlane(chg, TILE - 1);
    if ((threadIdx.x & 63) == 0) {
        wcol[threadIdx.x >> 6] = wc;
        ecol[threadIdx.x >> 6] = ec;
    }
}
__device__ inline void edge_columns_get(const uint32_t* wcol, const uint32_t* ecol,
                                        unsigned long long& w64, unsigned long long& e64) {
    w64 = 0;
    e64 = 0;
#pragma unroll
    for (int s = 0; s < TILE_THREADS / 64; ++s) {
        w64 |= (unsigned long long)wcol[s] << (STRIP * s);
        e64 |= (unsigned long long)ecol[s] << (STRIP * s);
    }
}

// Bits of the neighbour tiles whose halo contains local cell (lx, ly).
__device__ inline uint32_t border_bits(int lx, int ly) {
    uint32_t b = 0;
    const bool n = ly == 0, s = ly == TILE - 1, w = lx == 0, e = lx == TILE - 1;
    if (n) b |= 1u << 0;
    if (n && e) b |= 1u << 1;
    if (e) b |= 1u << 2;
    if (s && e) b |= 1u << 3;
    if (s) b |= 1u << 4;
    if (s && w) b |= 1u << 5;
    if (w) b |= 1u << 6;
    if (n && w) b |= 1u << 7;
    return b;
}

// Wave-wide OR of 8-bit masks and sum of values below 32, by ballots (no
// LDS permutes): the result in every lane.
__device__ inline uint32_t wave_or8(uint32_t bits) {
    uint32_t r = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d)
        if (__ballot((bits >> d) & 1u)) r |= 1u << d;
    return r;
}
__device__ inline uint32_t wave_sum5(uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) r += (uint32_t)__popcll(__ballot((v >> b) & 1u)) << b;
    return r;
}

// Set bits of the wave mask m in the lanes below this one.
__device__ inline uint32_t lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Rows of the neighbour in direction d (TDX/TDY order) that a change of
// this tile's border cells can unsettle: a halo row touches one row of the
// neighbour (its last row for the northern three, its first for the
// southern three); a halo column cell at row y touches rows y-1..y+1, with
// w64 / e64 the changed rows of columns 0 / TILE-1 (bit = row).
__device__ inline unsigned long long neighbour_rows(int d, unsigned long long w64,
                                                   unsigned long long e64) {
    if (d == 0 || d == 1 || d == 7) return 1ull << (TILE - 1);
    if (d == 3 || d == 4 || d == 5) return 1ull;
    const unsigned long long m = d == 2 ? e64 : w64;
    return m | (m << 1) | (m >> 1);
}

struct Ranked {
    // A list of `cap` entries filled from both ends: heavy tiles from the
    // front (the list's own count word), light ones from the back (count
    // *back), so the entries handed out first, by workgroup number, are the
    // heavy ones and a long solve overlaps the short ones instead of ending
    // the pass.
    const uint32_t* cost;
    uint32_t heavy;
    uint32_t* back;
    int cap;
};
__device__ inline int ranked_entry(const int* lst, int cap, uint32_t nfront, uint32_t i) {
    return i < nfront ? lst[i] : lst[cap - 1 - (int)(i - nfront)];
}
// Lists the neighbour tiles of tile (tx, ty) named by `bits` for the next
// pass, each at most once per epoch.  Called by all 64 lanes of one wave:
// lanes 0..7 take one direction each (their mark exchanges are in flight
// together), and one atomic reserves the list slots of the whole wave.
// With `rows` set, each neighbour's dirty-row mask for
// the next pass (two words at rows + 4 t + 2 rpar) also gets the rows the
// change unsettles; those words are read and cleared by the neighbour in
// that pass, and the other parity is the one being read in this pass.
// With `rank` set, the list is two-ended (see Ranked): a neighbour whose
// last solve took at least `rank->heavy` row steps goes to the front.
// With `open` set, a neighbour whose word there is 0 (it can no longer
// change) is not listed; the word is read with the mark exchange, so the
// check adds no round trip.  pushed_front / pushed_back (lane 0's, usually
// LDS words of the caller's thread 0) count the entries this call listed.
__device__ inline void activate_tiles(uint32_t bits, int tx, int ty, int ntx, int nty, int gbase,
                                      int* __restrict__ work_out, uint32_t* count_out,
                                      uint32_t* __restrict__ mark, uint32_t epoch,
                                      uint32_t* rows = nullptr, int rpar = 0,
                                      unsigned long long w64 = 0, unsigned long long e64 = 0,
                                      const Ranked* rank = nullptr,
                                      const uint32_t* open = nullptr,
                                      uint32_t* pushed_front = nullptr,
                                      uint32_t* pushed_back = nullptr) {
    const int lane = threadIdx.x & 63;
    bool push = false, heavy = true;
    int t = 0;
    if (lane < 8 && (bits & (1u << lane))) {
        const int nx = tx + TDX[lane], ny = ty + TDY[lane];
        if (nx >= 0 && ny >= 0 && nx < ntx && ny < nty) {
            t = gbase + ny * ntx + nx;
            uint32_t c = 0, o = 1;
            if (rank) c = rank->cost[t];  // in flight with the atomics below
            if (open) o = __hip_atomic_load(&open[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (rows) {
                const unsigned long long m = neighbour_rows(lane, w64, e64);
                uint32_t* p = rows + 4 * (size_t)t + 2 * rpar;
                if ((uint32_t)m) atomicOr(p, (uint32_t)m);
                if ((uint32_t)(m >> 32)) atomicOr(p + 1, (uint32_t)(m >> 32));
            }
            push = atomicExch(&mark[t], epoch) != epoch && o != 0u;
            if (rank) heavy = c >= rank->heavy;
        }
    }
    const unsigned long long mh = __ballot(push && heavy);
    const unsigned long long ml = __ballot(push && !heavy);
    if (mh) {
        uint32_t base = 0;
        if (lane == 0) {
            base = atomicAdd(count_out, (uint32_t)__popcll(mh));
            if (pushed_front) *pushed_front += (uint32_t)__popcll(mh);
        }
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        if (push && heavy) work_out[base + lanes_below(mh)] = t;
    }
    if (ml) {
        uint32_t base = 0;
        if (lane == 0) {
            base = atomicAdd(rank->back, (uint32_t)__popcll(ml));
            if (pushed_back) *pushed_back += (uint32_t)__popcll(ml);
        }
        base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        if (push && !heavy) work_out[rank->cap - 1 - (int)(base + lanes_below(ml))] = t;
    }
}

// Appends `item` for the lanes with `take` to the back of a two-ended list of
// `cap` entries (see Ranked): one atomic per wave.
__device__ inline void wave_append_back(bool take, int item, int* __restrict__ out,
                                        uint32_t* count, int cap) {
    const unsigned long long m = __ballot(take);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, first, 64);
    if (take) out[cap - 1 - (int)(base + lanes_below(m))] = item;
}

// Appends `item` for the lanes with `take` to a list: one atomic per wave.
__device__ inline void wave_append(bool take, int item, int* __restrict__ out, uint32_t* count) {
    const unsigned long long m = __ballot(take);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, first, 64);
    if (take) out[base + lanes_below(m)] = item;
}


}  // namespace cageo
