// Flow routing: FlowRoutingAlgorithm<T,U>::execute, FlowRoutingAlgorithm.h:
// 69-94; parity target FlowRoutingAlgorithm_CPU_OpenMP (D8 :53-151, flats
// :153-249 of FlowRoutingAlgorithmCPUOpenMP.h).
//
// D8: edge cells point out of the grid; cells at 0 are stream cells with no
// direction; otherwise the steepest of N, NE, E, SE, S, SW, W, NW with
// drop = z_c - z_n, diagonal drops divided by 1.41421356237309505f, strict
// '>' so the first of equal drops wins.  A cell with no positive drop is
// flat.
//
// Flats: the reference resolves them in rounds k = -1, 1, 2, ...: a flat c
// takes every neighbour n (same order, the last one wins) with
// flatDist[n] == k and z_n <= z_c.  That is a breadth-first distance
//   d(c) = 1 + min { d(n) : n an 8-neighbour, z_n <= z_c,
//                            n has a direction (d = 0) or is a flat },
// followed by direction = last n in order with d(n) = d(c) - 1 and
// z_n <= z_c.  d is computed by monotone min-relaxation over 64x64 LDS
// tiles (exact row solves by min-plus wave scans) inside one persistent
// kernel, with the same worklist protocol as the pit fill.
#include <vector>

#include "tiles.hpp"

namespace cageo {

namespace {

constexpr int D_INF = 0x3fffffff;
constexpr int D_STREAM = -2;
constexpr int D_OUT = -3;

__device__ constexpr int FDX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ constexpr int FDY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};

struct FrShared {
    int2 dz[HALO_W * HALO_W];  // x: distance, y: bits of z
    uint32_t rowdirty[TILE];   // strip edge rows marked by the neighbouring strips
    uint32_t wcol[TILE_THREADS / 64], ecol[TILE_THREADS / 64];  // changed rows of columns 0, TILE-1
    uint32_t bits;
    uint32_t any;
    uint32_t pushed;  // tiles this workgroup listed in the pass (thread 0's)
};

// Min-plus map v -> min(a, v + k) on distances saturated at D_INF (k = D_INF
// means the map ignores v).  Replaces (a, k) by "self after prev".
__device__ inline int sat_add(int x, int y) {
    const int s = x + y;  // both <= D_INF < 2^30: no overflow
    return s < D_INF ? s : D_INF;
}
__device__ inline void minplus_after(int& a, int& k, int pa, int pk) {
    a = min(a, sat_add(pa, k));
    k = sat_add(pk, k);
}
// Inclusive min-plus scans towards higher / lower lanes, their lane moves'
// destinations carried across row steps (ScanCarry).
typedef ScanCarry<int> MinplusCarry;
__device__ inline void minplus_carry_init(MinplusCarry& c) { scan_carry_init(c, (int)D_INF, 0); }
__device__ inline void scan_minplus_up(int& a, int& k, MinplusCarry& c) {
    scan_up_carried(a, k, c, minplus_after);
}
__device__ inline void scan_minplus_down(int& a, int& k, MinplusCarry& c) {
    scan_down_carried(a, k, c, (int)D_INF, 0, minplus_after);
}

// One raster of a launch.
struct FrRaster {
    const float* z;
    int8_t* fd;
    int32_t* dist;
    int W, H, ntx, nty;
};

// A launch over one raster (`one`, rs null) or over the nr rasters of a
// table, whose 64x64 blocks are numbered globally (bmap).
struct FrArgs {
    FrRaster one;
    const FrRaster* __restrict__ rs;
    const int2* __restrict__ bmap;  // per block: (raster, block within it)
    int nr, total;
    int* work;      // 3 lists of `total` entries
    uint32_t* mark; // per-block epoch stamps
    uint32_t* rows; // per-block dirty-row masks: 2 pass parities x 64 bits
    uint32_t* ctrl;
    GridBar gb;     // grid barrier
    int max_passes;
};

struct FrBlock {
    FrRaster R;
    int lb, gbase;
};
template <bool MULTI>
__device__ inline FrBlock fr_block(const FrArgs& a, int g) {
    FrBlock b;
    g = __builtin_amdgcn_readfirstlane(g);
    if constexpr (!MULTI) {
        b.R = a.one;
        b.gbase = 0;
    } else {
        const int2 m = a.bmap[g];
        const int r = __builtin_amdgcn_readfirstlane(m.x);
        b.gbase = g - __builtin_amdgcn_readfirstlane(m.y);
        b.R = a.rs[r];
    }
    b.lb = g - b.gbase;
    return b;
}


__device__ inline uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// D8 of one cell (FlowRoutingAlgorithmCPUOpenMP.h:60-150) at grid position
// (x, y) from a halo tile of elevations `zt` (rows `stride` apart; the cell
// at index c).  Returns the initial distance: 0 has a direction, D_STREAM,
// or D_INF (flat); (ox, oy) the direction.  Edge cells point out of the grid
// and read no neighbour, so halo cells outside the grid are never used.
__device__ inline int32_t d8_at(const float* zt, int stride, int c, int x, int y, int W, int H,
                                int8_t& ox, int8_t& oy) {
    int dxi = 0, dyi = 0;
    if (x == 0) dxi = -1;
    else if (x == W - 1) dxi = 1;
    if (y == 0) dyi = -1;
    else if (y == H - 1) dyi = 1;
    ox = 0;
    oy = 0;
    if (dxi != 0 || dyi != 0) {
        ox = (int8_t)dxi;
        oy = (int8_t)dyi;
        return 0;
    }
    const float zc = zt[c];
    if (zc == 0.0f) return D_STREAM;
    // the direction is carried as its offsets, each a constant of the
    // unrolled loop (indexing the tables by a variable would load them from
    // memory)
    float steepest = 0.0f;
    int32_t d = D_INF;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float s = zc - zt[c + FDY[k] * stride + FDX[k]];
        if (s > 0.0f) {
            if (k & 1) s = s / 1.41421356237309505f;
            if (s > steepest) {
                steepest = s;
                ox = (int8_t)FDX[k];
                oy = (int8_t)FDY[k];
                d = 0;
            }
        }
    }
    return d;
}

__device__ void bfs_solve(FrShared& S, const FrArgs& args, const FrBlock& b, int tx, int ty,
                          unsigned long long dirty, int* work_out, uint32_t* count_out,
                          uint32_t epoch, int pass, bool write_all);

// A listed tile of a flat pass: its halo tile of distances and elevations
// loaded, then solved (bfs_solve).
__device__ void bfs_tile(FrShared& S, const FrArgs& args, const FrBlock& b, int* work_out,
                         uint32_t* count_out, uint32_t epoch, int pass) {
    const FrRaster& a = b.R;
    const unsigned long long dirty = tile_dirty_rows(args.rows, b.gbase + b.lb, pass);
    if (dirty == 0) return;  // uniform
    const int tx = b.lb % a.ntx, ty = b.lb / a.ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    // halo tile in two batches of independent loads (each batch's loads are
    // in flight together; one batch would take the registers of the fourth
    // workgroup a CU can hold)
    {
        constexpr int N = HALO_W * HALO_W;
        constexpr int ITER = (N + TILE_THREADS - 1) / TILE_THREADS;
        constexpr int CHUNK = (ITER + 1) / 2;
#pragma unroll
        for (int c0 = 0; c0 < ITER; c0 += CHUNK) {
            int2 v[CHUNK];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const int k = threadIdx.x + (c0 + j) * TILE_THREADS;
                v[j] = make_int2(D_OUT, 0);
                if (c0 + j < ITER && k < N) {
                    const int hy = k / HALO_W, hx = k - hy * HALO_W;
                    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
                    if (gx >= 0 && gy >= 0 && gx < a.W && gy < a.H) {
                        const size_t gi = (size_t)gy * a.W + gx;
                        v[j] = make_int2(a.dist[gi], __float_as_int(a.z[gi]));
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const int k = threadIdx.x + (c0 + j) * TILE_THREADS;
                if (c0 + j < ITER && k < N) S.dz[k] = v[j];
            }
        }
    }
    if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
    if (threadIdx.x == 0) S.bits = 0;
    __syncthreads();
    clear_dirty_rows(args.rows, b.gbase + b.lb, pass);
    bfs_solve(S, args, b, tx, ty, dirty, work_out, count_out, epoch, pass, false);
}

// The flat distances of a tile whose distances and elevations (with the
// halo) are in LDS, S.rowdirty and S.bits clear: row solves over the rows of
// `dirty` first, then every row a change reaches; the changed distances
// written back (with write_all every distance of the block's cells), and
// the neighbour tiles a changed border cell concerns listed for pass
// `pass` + 1.
__device__ void bfs_solve(FrShared& S, const FrArgs& args, const FrBlock& b, int tx, int ty,
                          unsigned long long dirty, int* work_out, uint32_t* count_out,
                          uint32_t epoch, int pass, bool write_all) {
    const FrRaster& a = b.R;
    const int x0 = tx * TILE, y0 = ty * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t flat = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int li = (strip * STRIP + k + 1) * HALO_W + lx + 1;
        if (S.dz[li].x >= 1) flat |= 1u << k;
    }
    // Row step: a flat cell takes d = 1 + min d(n) over eligible neighbours
    // (d(n) >= 0, z_n <= z_c).  Rows above and below and the halo cells
    // enter as a constant; along the row each flat cell is the map
    // v -> min(own, v + 1) when the row neighbour it reads is not higher,
    // cells with a direction relay their 0, stream and outside cells relay
    // nothing.  Scans from both sides give the row's exact fixed point.
    uint32_t changed_cells = 0;
    uint32_t rd = strip_rows(dirty, strip);
    MinplusCarry mc;
    minplus_carry_init(mc);
    for (int sweep = 0;; ++sweep) {
        int changed = 0;
        const bool upward = sweep_upward(S.rowdirty, rd, strip, sweep);
#pragma unroll
        for (int kk = 0; kk < STRIP; ++kk) {
            const int k = upward ? (STRIP - 1 - kk) : kk;
            const int ly = strip * STRIP + k;
            if (!row_due(S.rowdirty, rd, k, ly)) continue;
            const bool isflat = (flat >> k) & 1u;
            if (__ballot(isflat) == 0) continue;  // no flat cell: nothing in the row can change
            const int li = (ly + 1) * HALO_W + lx + 1;
            // the step's reads at once: the cell and its eight neighbours
            const int2 c = S.dz[li];
            int2 nb[8];
#pragma unroll
            for (int d = 0; d < 8; ++d) nb[d] = S.dz[li + FDY[d] * HALO_W + FDX[d]];
            const float zc = __int_as_float(c.y);
            const float zl = lane_prev(zc), zrt = lane_next(zc);
            int a, kl = D_INF, kr = D_INF;
            if (isflat) {
                a = c.x;
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    // rows above and below; the halo columns for the end lanes
                    const bool in_row = FDY[d] == 0;
                    if (in_row && !((lx == 0 && FDX[d] < 0) || (lx == TILE - 1 && FDX[d] > 0)))
                        continue;
                    if (nb[d].x >= 0 && __int_as_float(nb[d].y) <= zc) a = min(a, nb[d].x + 1);
                }
                if (lx > 0 && zl <= zc) kl = 1;
                if (lx < TILE - 1 && zrt <= zc) kr = 1;
            } else {
                a = c.x >= 0 ? c.x : D_INF;  // a direction relays 0
            }
            int al = a, ar = a;
            scan_minplus_up(al, kl, mc);
            scan_minplus_down(ar, kr, mc);
            bool ch = false;
            int nd = D_INF;
            if (isflat) {
                nd = min(al, ar);
                if (nd < c.x) {
                    S.dz[li].x = nd;
                    changed_cells |= 1u << k;
                    ch = true;
                }
            }
            if (__ballot(ch)) {
                changed = 1;
                // The rows above and below are due again only where a cell
                // can take the new distance: a flat (d >= 1) not lower than
                // the changed cell whose distance exceeds the new one plus
                // one (distances only decrease, so a neighbour strip's row
                // read before that strip lowered it errs towards marking).
                // Cells of the halo columns belong to the next tiles.
                bool up = false, dn = false;
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    if (FDY[d] == 0) continue;
                    if ((lx == 0 && FDX[d] < 0) || (lx == TILE - 1 && FDX[d] > 0)) continue;
                    const bool t = ch && nb[d].x >= 1 && nb[d].x > nd + 1 &&
                                   __int_as_float(nb[d].y) >= zc;
                    if (FDY[d] < 0) up |= t;
                    else dn |= t;
                }
                row_mark_dirs(S.rowdirty, rd, k, ly, __ballot(up) != 0, __ballot(dn) != 0);
            }
        }
        if (!__syncthreads_or(changed)) break;
    }
    uint32_t bits = 0, seen = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const bool in = x0 + lx < a.W && y0 + ly < a.H;
        if (in && (write_all || (changed_cells & (1u << k))))
            a.dist[(size_t)(y0 + ly) * a.W + x0 + lx] = S.dz[(ly + 1) * HALO_W + lx + 1].x;
    }
    // A changed border cell concerns a neighbour tile only if one of its
    // cells in this tile's halo can take the new distance: a flat (d >= 1)
    // not lower than the cell whose distance, as loaded (distances only
    // decrease), exceeds the new one plus one.  Only such cells activate
    // the tile they lie in.  A loop over the few border cells (not
    // unrolled: the row is a variable).
    {
        const uint32_t edge = (lx == 0 || lx == TILE - 1)
                                  ? 0xffffu
                                  : ((strip == 0 ? 1u : 0u) |
                                     (strip == TILE_THREADS / 64 - 1 ? 1u << (STRIP - 1) : 0u));
        uint32_t cand = changed_cells & edge;
        while (cand) {
            const int k = __builtin_ctz(cand);
            cand &= cand - 1;
            const int ly = strip * STRIP + k;
            const int li = (ly + 1) * HALO_W + lx + 1;
            const int2 cv = S.dz[li];
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                const int nx = lx + FDX[d], ny = ly + FDY[d];
                if (nx >= 0 && ny >= 0 && nx < TILE && ny < TILE) continue;
                const int2 nv = S.dz[li + FDY[d] * HALO_W + FDX[d]];
                if (nv.x >= 1 && nv.x > cv.x + 1 &&
                    __int_as_float(nv.y) >= __int_as_float(cv.y)) {
                    bits |= 1u << tile_dir(nx < 0 ? -1 : (nx >= TILE ? 1 : 0),
                                           ny < 0 ? -1 : (ny >= TILE ? 1 : 0));
                    seen |= 1u << k;
                }
            }
        }
    }
    bits = wave_or8(bits);
    edge_columns_put(S.wcol, S.ecol, seen);
    if ((threadIdx.x & 63) == 0 && bits) atomicOr(&S.bits, bits);
    __syncthreads();
    if (threadIdx.x < 64 && S.bits) {  // uniform within the first wave
        unsigned long long w64, e64;
        edge_columns_get(S.wcol, S.ecol, w64, e64);
        activate_tiles(S.bits, tx, ty, a.ntx, a.nty, b.gbase, work_out, count_out, args.mark,
                       epoch, args.rows, (pass + 1) & 1, w64, e64, nullptr, nullptr,
                       &S.pushed);
    }
    __syncthreads();
}

// The epoch stamp of pass 0's listings (the D8 kernel's); pass p lists with
// FR_EPOCH0 + p.  k_fr_reset clears every stamp to 0 before a launch, so the
// epochs can start over each launch and never come round to 0.
constexpr uint32_t FR_EPOCH0 = 1u;

// Before a launch's D8 kernel: every block's epoch stamp and dirty rows
// cleared (an abandoned launch may have left some, and the D8 kernel's
// blocks list each other, so none can clear its own there) and the worklist
// counts zeroed.
__global__ void k_fr_reset(FrArgs a) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < a.total; g += gridDim.x * blockDim.x) {
        a.mark[g] = 0u;
        reinterpret_cast<uint4*>(a.rows)[g] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x < 3) a.ctrl[CTL_COUNT + threadIdx.x] = 0u;
        if (threadIdx.x == 3) a.ctrl[CTL_FR_DONE] = 0u;
    }
}

// D8 and the first flat pass, one workgroup per block.  The block's
// elevations with a two-cell halo are staged in LDS, D8 is worked out for
// the 66 x 66 cells of its distance tile (its own and the ring around it:
// the ring's are the initial distances its neighbours start from), and the
// first pass solves the tile against that ring as the persistent kernel's
// first pass did against the distances the D8 kernel had written; then
// the block's distances and the directions of its cells go out, and the
// tiles its changed border cells concern are listed for pass 1 (list 1,
// epoch FR_EPOCH0).  The persistent kernel goes on from pass 1.
template <bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_fr_d8(FrArgs a) {
    __shared__ FrShared S;
    const FrBlock b = fr_block<MULTI>(a, blockIdx.x);
    const FrRaster& R = b.R;
    const int tx = b.lb % R.ntx, ty = b.lb / R.ntx;
    const int x0 = tx * TILE, y0 = ty * TILE;
    constexpr int ZW = HALO_W + 2, ZN = ZW * ZW;  // elevations with two halo cells
    constexpr int N = HALO_W * HALO_W;            // the distance tile
    static_assert(ZN * sizeof(float) <= sizeof(S.dz), "the elevation halo fits the distance tile");
    float* zt = reinterpret_cast<float*>(S.dz);
    {
        constexpr int IT = (ZN + TILE_THREADS - 1) / TILE_THREADS;
        float v[IT];
#pragma unroll
        for (int j = 0; j < IT; ++j) {  // all loads first (a cell outside loads cell 0)
            const int k = threadIdx.x + j * TILE_THREADS;
            const int hy = k / ZW, hx = k - hy * ZW;
            const int gx = x0 - 2 + hx, gy = y0 - 2 + hy;
            const bool in = k < ZN && gx >= 0 && gy >= 0 && gx < R.W && gy < R.H;
            v[j] = R.z[in ? (size_t)gy * R.W + gx : 0];
            if (!in) v[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < IT; ++j) {
            const int k = threadIdx.x + j * TILE_THREADS;
            if (k < ZN) zt[k] = v[j];
        }
    }
    __syncthreads();
    constexpr int ITER = (N + TILE_THREADS - 1) / TILE_THREADS;
    int2 dzv[ITER];
#pragma unroll
    for (int j = 0; j < ITER; ++j) {
        const int k = threadIdx.x + j * TILE_THREADS;
        const int hy = k / HALO_W, hx = k - hy * HALO_W;
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
        dzv[j] = make_int2(D_OUT, 0);
        if (k < N && gx >= 0 && gy >= 0 && gx < R.W && gy < R.H) {
            const int c = (hy + 1) * ZW + hx + 1;
            int8_t ox, oy;
            const int32_t d = d8_at(zt, ZW, c, gx, gy, R.W, R.H, ox, oy);
            dzv[j] = make_int2(d, __float_as_int(zt[c]));
            if (hx >= 1 && hy >= 1 && hx <= TILE && hy <= TILE)  // the block's own cell
                reinterpret_cast<char2*>(R.fd)[(size_t)gy * R.W + gx] = make_char2(ox, oy);
        }
    }
    __syncthreads();  // every read of the elevation halo is done
#pragma unroll
    for (int j = 0; j < ITER; ++j) {
        const int k = threadIdx.x + j * TILE_THREADS;
        if (k < N) S.dz[k] = dzv[j];
    }
    if (threadIdx.x < TILE) S.rowdirty[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
        S.bits = 0;
        S.pushed = 0;
    }
    __syncthreads();
    bfs_solve(S, a, b, tx, ty, ~0ull, a.work + (size_t)a.total, &a.ctrl[CTL_COUNT + 1],
              FR_EPOCH0, 0, true);
}

// The flat passes, after k_fr_d8 (k_fr_pick follows).
// MULTI: a table of rasters (else the one raster of the arguments).
// Four waves a SIMD: the LDS of a flat tile allows four workgroups per CU.
template <bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_fr_persistent(FrArgs a) {
    __shared__ FrShared S;
    if (grid_aborted(a.gb)) return;
    uint32_t epoch = FR_EPOCH0;

    int li = 1;  // pass 0 ran in k_fr_d8 and filled list 1
    bool converged = false;
    int pass = 1;
    // after a pass the next list's count comes with the barrier (the sum of
    // what each workgroup listed) where the blocks are fewer than
    // BAR_SUM_LIMIT
    const bool sums = (uint64_t)a.total < BAR_SUM_LIMIT;
    uint32_t next_count = 0;
    for (; pass < a.max_passes; ++pass) {
        const int cin = li, cout = (li + 1) % 3, cclr = (li + 2) % 3;
        const uint32_t count = pass > 1 && sums ? next_count : ld(&a.ctrl[CTL_COUNT + cin]);
        if (count == 0) {
            converged = true;
            break;
        }
        ++epoch;
        if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl[CTL_COUNT + cclr] = 0;
        const int* win = a.work + (size_t)cin * a.total;
        int* wout = a.work + (size_t)cout * a.total;
        if (threadIdx.x == 0) S.pushed = 0u;
        for (uint32_t i = blockIdx.x; i < count; i += gridDim.x)
            bfs_tile(S, a, fr_block<MULTI>(a, win[i]), wout, &a.ctrl[CTL_COUNT + cout], epoch, pass);
        uint32_t unused = 0;
        if (!grid_sync_sum(a.gb, 0, sums ? S.pushed : 0u, 0u, next_count, unused)) return;
        li = cout;
    }
    // The words written below are not read by the pass loop, so no barrier
    // is needed between the last count read and them.
    if (!converged) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            a.ctrl[CTL_STATUS] = (uint32_t)(-CAGEO_ERR_NOT_CONVERGED);
        return;
    }

    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl[CTL_FR_PASSES] = (uint32_t)pass;  // read by the host only
        a.ctrl[CTL_FR_DONE] = 1u;               // k_fr_pick may go on
    }
}

// Directions of the resolved flats (FlowRoutingAlgorithmCPUOpenMP.h:
// 153-249): the last neighbour in order with d(n) = d(c) - 1 and
// z_n <= z_c.  One workgroup per block: its own distances first (a block
// without a resolved flat is done), then distances and elevations with a
// one-cell halo staged in LDS in one batch of loads; a halo cell outside the
// grid reads D_OUT, which no d(c) - 1 >= 0 equals.  Runs only after flat
// passes that converged (k_fr_persistent sets CTL_FR_DONE).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_fr_pick(FrArgs a) {
    __shared__ int2 dz[HALO_W * HALO_W];  // x: distance, y: bits of z
    if (ld(&a.ctrl[CTL_FR_DONE]) == 0u) return;
    const FrBlock b = fr_block<MULTI>(a, blockIdx.x);
    const FrRaster& R = b.R;
    const int x0 = (b.lb % R.ntx) * TILE, y0 = (b.lb / R.ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = x0 + lx;
    uint32_t flat = 0;
    {
        int32_t dv[STRIP];
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            const int y = y0 + strip * STRIP + k;
            const bool in = x < R.W && y < R.H;
            dv[k] = R.dist[in ? (size_t)y * R.W + x : 0];
            if (!in) dv[k] = 0;
        }
#pragma unroll
        for (int k = 0; k < STRIP; ++k)
            if (dv[k] >= 1 && dv[k] < D_INF) flat |= 1u << k;
    }
    if (!__syncthreads_or(flat != 0u)) return;
    {
        constexpr int N = HALO_W * HALO_W;
        constexpr int ITER = (N + TILE_THREADS - 1) / TILE_THREADS;
        int2 v[ITER];
#pragma unroll
        for (int j = 0; j < ITER; ++j) {
            const int k = threadIdx.x + j * TILE_THREADS;
            const int hy = k / HALO_W, hx = k - hy * HALO_W;
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
            const bool in = k < N && gx >= 0 && gy >= 0 && gx < R.W && gy < R.H;
            const size_t gi = in ? (size_t)gy * R.W + gx : 0;
            v[j] = make_int2(R.dist[gi], __float_as_int(R.z[gi]));
            if (!in) v[j].x = D_OUT;
        }
#pragma unroll
        for (int j = 0; j < ITER; ++j) {
            const int k = threadIdx.x + j * TILE_THREADS;
            if (k < N) dz[k] = v[j];
        }
    }
    __syncthreads();
    while (flat) {
        const int k = __builtin_ctz(flat);
        flat &= flat - 1;
        const int ly = strip * STRIP + k;
        const int li = (ly + 1) * HALO_W + lx + 1;
        const int2 c = dz[li];
        const float zc = __int_as_float(c.y);
        bool found = false;
        char2 dir = make_char2(0, 0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int2 n = dz[li + FDY[q] * HALO_W + FDX[q]];
            if (n.x == c.x - 1 && __int_as_float(n.y) <= zc) {  // the last match
                found = true;
                dir = make_char2((int8_t)FDX[q], (int8_t)FDY[q]);
            }
        }
        if (found) reinterpret_cast<char2*>(R.fd)[(size_t)(y0 + ly) * R.W + x] = dir;
    }
}

FrRaster fr_raster(Context* c, const float* z, int8_t* fd) {
    FrRaster R;
    R.z = z;
    R.fd = fd;
    R.dist = c->d_dist;
    R.W = c->W;
    R.H = c->H;
    R.ntx = c->ntx;
    R.nty = c->nty;
    return R;
}

}  // namespace

// Co-resident workgroups of the persistent kernel on this device.
int flow_routing_grid_cap(Context* c, int* cap) {
    int per_cu = 0;
    CAGEO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_fr_persistent<false>), TILE_THREADS, 0));
    hipDeviceProp_t prop;
    CAGEO_HIP_TRY(hipGetDeviceProperties(&prop, c->device));
    if (per_cu < 1) return CAGEO_ERR_HIP;
    *cap = per_cu * prop.multiProcessorCount;
    return CAGEO_OK;
}

int flow_routing(Context* c, const float* z, int8_t* fd) {
    if (c->fr_grid <= 0) CAGEO_TRY(flow_routing_grid_cap(c, &c->fr_grid));
    FrArgs a;
    a.one = fr_raster(c, z, fd);
    a.rs = nullptr;
    a.bmap = nullptr;
    a.nr = 1;
    a.total = c->ntiles;
    a.work = c->d_work3;
    a.mark = c->d_mark;
    a.rows = c->d_rows;
    a.ctrl = c->d_ctrl;
    a.gb = context_gridbar(c);
    a.max_passes = 1 << 20;
    hipLaunchKernelGGL(k_fr_reset, dim3((a.total + 255) / 256), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_fr_d8<false>, dim3(a.total), dim3(TILE_THREADS), 0, c->stream, a);
    // ordinary launch within the occupancy cap (see pit_fill_quantized)
    hipLaunchKernelGGL(k_fr_persistent<false>, dim3(c->fr_grid), dim3(TILE_THREADS), 0, c->stream,
                       a);
    hipLaunchKernelGGL(k_fr_pick<false>, dim3(a.total), dim3(TILE_THREADS), 0, c->stream, a);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

// Raster table of a batch: each context's d_w -> d_fd.
int flow_routing_tables(MultiState* m, Context* const* ctxs) {
    std::vector<FrRaster> t(m->nr);
    for (int r = 0; r < m->nr; ++r) t[r] = fr_raster(ctxs[r], ctxs[r]->d_w, ctxs[r]->d_fd);
    CAGEO_HIP_TRY(hipMalloc(&m->d_fr, t.size() * sizeof(FrRaster)));
    CAGEO_HIP_TRY(hipMemcpy(m->d_fr, t.data(), t.size() * sizeof(FrRaster), hipMemcpyHostToDevice));
    {
        int per_cu = 0;
        CAGEO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(k_fr_persistent<true>), TILE_THREADS, 0));
        hipDeviceProp_t prop;
        CAGEO_HIP_TRY(hipGetDeviceProperties(&prop, ctxs[0]->device));
        if (per_cu < 1) return CAGEO_ERR_HIP;
        m->fr_cap = per_cu * prop.multiProcessorCount;
    }
    if (m->fr_grid <= 0) m->fr_grid = m->fr_cap;
    return CAGEO_OK;
}

int flow_routing_multi(MultiState* m, hipStream_t s) {
    FrArgs a;
    a.one = FrRaster{};
    a.rs = static_cast<const FrRaster*>(m->d_fr);
    a.bmap = m->d_bmap;
    a.nr = m->nr;
    a.total = m->total;
    a.work = m->d_work;
    a.mark = m->d_mark_fr;
    a.rows = m->d_rows;
    a.ctrl = m->d_ctrl;
    a.gb = multi_gridbar(m);
    a.max_passes = 1 << 20;
    hipLaunchKernelGGL(k_fr_reset, dim3((a.total + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_fr_d8<true>, dim3(a.total), dim3(TILE_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_fr_persistent<true>, dim3(m->fr_grid), dim3(TILE_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_fr_pick<true>, dim3(a.total), dim3(TILE_THREADS), 0, s, a);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
