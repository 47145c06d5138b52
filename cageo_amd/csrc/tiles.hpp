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

// ---- exact row solves of the LDS tile sweeps -------------------------------
// A wave owns a strip of STRIP tile rows, one lane per column.  Each step
// solves one row exactly with its two neighbour rows held at their current
// values: along the row the relaxation is a chain of per-cell monotone maps,
// composed for all 64 columns at once by wave scans (6 shuffle steps in
// each direction), so information crosses the whole row in one step rather
// than one column per sweep.  Rows are visited in alternating order (Gauss-
// Seidel down, then up); a row whose cells changed marks its neighbour rows
// for another solve, and a tile is done after a sweep without changes.

// Clamp map v -> min(max(v, lo), hi) with lo <= hi.  Replaces (lo, hi) by
// the composition "self after prev".  Compositions of clamps are clamps.
__device__ inline void clamp_after(float& lo, float& hi, float plo, float phi) {
    const float nlo = fminf(fmaxf(plo, lo), hi);
    const float nhi = fminf(fmaxf(phi, lo), hi);
    lo = nlo;
    hi = nhi;
}

// Inclusive scan towards higher lanes: lane x ends with f_x o ... o f_0.
__device__ inline void scan_clamp_up(float& lo, float& hi) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float plo = __shfl_up(lo, d, 64), phi = __shfl_up(hi, d, 64);
        if (lane >= d) clamp_after(lo, hi, plo, phi);
    }
}

// Inclusive scan towards lower lanes: lane x ends with f_x o ... o f_63.
__device__ inline void scan_clamp_down(float& lo, float& hi) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float plo = __shfl_down(lo, d, 64), phi = __shfl_down(hi, d, 64);
        if (lane + d < 64) clamp_after(lo, hi, plo, phi);
    }
}

// Row flags of the sweeps: rows ly-1 and ly+1, and ly itself when `self`,
// need another solve.  Plain stores of 1 (the flags only ever go 0 -> 1
// between the owner's read-and-clear steps).
__device__ inline void mark_rows(uint32_t* rd, int ly, bool self) {
    if ((threadIdx.x & 63) == 0) {
        if (ly > 0) rd[ly - 1] = 1u;
        if (ly < TILE - 1) rd[ly + 1] = 1u;
        if (self) rd[ly] = 1u;
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

// Called by one thread per workgroup.
__device__ inline void activate_tiles(uint32_t bits, int tx, int ty, int ntx, int nty,
                                      int* __restrict__ work_out, uint32_t* count_out,
                                      uint32_t* __restrict__ mark, uint32_t epoch) {
    for (int d = 0; d < 8; ++d) {
        if (!(bits & (1u << d))) continue;
        int nx = tx + TDX[d], ny = ty + TDY[d];
        if (nx < 0 || ny < 0 || nx >= ntx || ny >= nty) continue;
        int t = ny * ntx + nx;
        if (atomicExch(&mark[t], epoch) != epoch) {
            uint32_t slot = atomicAdd(count_out, 1u);
            work_out[slot] = t;
        }
    }
}

// Block-wide OR of per-thread border bits; returns the result in all threads.
__device__ inline uint32_t block_or_bits(uint32_t bits) {
    __shared__ uint32_t s_bits;
    if (threadIdx.x == 0) s_bits = 0;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) bits |= (uint32_t)__shfl_xor((int)bits, off);
    if ((threadIdx.x & 63) == 0 && bits) atomicOr(&s_bits, bits);
    __syncthreads();
    return s_bits;
}

// Host side: run passes until no tile is listed.  `launch(work_in, count,
// work_out, count_slot, epoch)` enqueues one pass.  `count` tiles are in
// d_work[buf] on entry.
template <class Launch>
int run_tile_passes(Context* c, int buf, uint32_t count, int max_passes, Launch launch,
                    int* passes_out) {
    int passes = 0;
    while (count > 0) {
        if (passes >= max_passes) return CAGEO_ERR_NOT_CONVERGED;
        const int out = 1 - buf;
        const uint32_t epoch = next_epoch(c);
        CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_COUNT0 + out, 0, sizeof(uint32_t), c->stream));
        CAGEO_TRY(launch(c->d_work[buf], count, c->d_work[out], c->d_scal + SC_COUNT0 + out, epoch));
        CAGEO_TRY(read_scalars(c));
        count = c->h_scal[SC_COUNT0 + out];
        buf = out;
        ++passes;
    }
    if (passes_out) *passes_out += passes;
    return CAGEO_OK;
}

}  // namespace cageo
