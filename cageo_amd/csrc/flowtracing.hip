// Flow tracing: FlowTracingAlgorithm<T,U,V>::execute, FlowTracingAlgorithm.h:
// 52-69; parity target FlowTracingAlgorithmCPUOpenMP.h:34-134.
//
// The reference walks every untraced cell along its directions until it
// leaves the grid (OUTSIDE), steps onto a stream cell (label = -stream as
// int) or onto an already labelled cell (copies it); stream cells themselves
// are LAKE; a walk longer than 100 001 steps (a cell with no direction, or a
// cycle) leaves its path at 0.  A cell's label is therefore the label of its
// path's terminal, computed here by pointer jumping in two levels:
//   1. per 64x64 tile, in LDS: every cell's pointer jumps along the in-tile
//      part of its path (12 doubling rounds cover 4096 cells) and ends as a
//      terminal label, "stuck" (its path cycles inside the tile, e.g. a cell
//      without a direction), or an exit: the first cell outside the tile.
//      Cells that end in their tile get their label at once.  An exit lands
//      on a border cell of the neighbouring tile, so the tile also writes the
//      entries of its own 252 border cells into a compact border table
//      (256 slots per tile), and each exit cell keeps the border slot it
//      exits to and its step count;
//   2. over the border table only: exits jump through the entries of the
//      slots they name, in place, until no exit is left; a chain still moving
//      after ceil(log2 slots)+2 rounds is a cycle through several tiles
//      (stuck).  In place is safe because every entry of a slot, old or
//      rewritten, names a slot on that cell's own path with the step count to
//      it, so a hop through either is correct, and a rewritten one is further
//      along;
//   3. every exit cell takes the label its border slot resolved to.
// Entries are 64 bits: kind (2 bits) | steps (30 bits) | value (32 bits), so
// the step count to the terminal is known and cells whose terminal is more
// than 100 001 steps away (where the reference's result depends on its
// raster order) are counted.  performConnectFlowPath (:115-134) then maps a
// label >= 0 (only 0 occurs: labels are -stream <= 0 or the negative
// constants) to label[0] when that is negative; it is applied in step 3.
#include <vector>

#include "common.hpp"

namespace cageo {

namespace {

constexpr uint32_t K_LOCAL = 0, K_RESOLVED = 1, K_EXIT = 2, K_STUCK = 3;
constexpr uint64_t STEP_MASK = (1ull << 30) - 1;
// Label of an exit cell between steps 1 and 3 (labels are -stream ids, at
// least -2^31 + 1, or the negative constants, or 0).
constexpr int32_t EXIT_MARK = INT32_MIN;
constexpr int BORDER_SLOTS = 256;

// Slot of tile cell (lx, ly) in its tile's border table: row 0 -> 0..63,
// row 63 -> 64..127, column 0 (rows 1..62) -> 129..190, column 63 (rows
// 1..62) -> 193..254; -1 inside.  Slots 128, 191, 192 and 255 stay unused.
__device__ inline int border_slot(int lx, int ly) {
    if (ly == 0) return lx;
    if (ly == TILE - 1) return TILE + lx;
    if (lx == 0) return 2 * TILE + ly;
    if (lx == TILE - 1) return 3 * TILE + ly;
    return -1;
}

__device__ inline uint64_t pack(uint32_t kind, uint32_t steps, int32_t value) {
    return ((uint64_t)kind << 62) | ((uint64_t)(steps & STEP_MASK) << 32) |
           (uint64_t)(uint32_t)value;
}
__device__ inline uint32_t kind_of(uint64_t e) { return (uint32_t)(e >> 62); }
__device__ inline uint32_t steps_of(uint64_t e) { return (uint32_t)((e >> 32) & STEP_MASK); }
__device__ inline int32_t value_of(uint64_t e) { return (int32_t)(uint32_t)e; }

// Follow entry e one hop through entry f (the entry of e's target).
__device__ inline uint64_t hop(uint64_t e, uint64_t f) {
    uint32_t s = steps_of(e) + steps_of(f);
    if (s > STEP_MASK) s = STEP_MASK;
    const uint32_t k = kind_of(f);
    if (k == K_STUCK) return pack(K_STUCK, 0, 0);
    return pack(k, s, value_of(f));
}

// Inlined: as a call its frame spilled to scratch and its registers set the
// kernel's occupancy.
__device__ __forceinline__ void trace_tile(uint64_t* L, const int8_t* __restrict__ fd,
                                           const uint8_t* __restrict__ sf,
                                           const int32_t* __restrict__ slabel,
                                           uint64_t* __restrict__ ent, int32_t* __restrict__ lab,
                                           uint64_t* __restrict__ bent, int W, int H, int ntx,
                                           int t) {
    const int x0 = (t % ntx) * TILE, y0 = (t / ntx) * TILE;
    const int lx = threadIdx.x & 63, strip = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gx = x0 + lx;

    // The strip's cells' flags and directions are loaded together, then
    // their targets' flags and labels: two dependent rounds of loads for the
    // tile (no load waits on a test; a target outside the grid reads the
    // cell itself, unused).
    uint32_t local = 0;
    constexpr int HALF = STRIP;  // rows per round of loads: all of the strip
#pragma unroll
    for (int k0 = 0; k0 < STRIP; k0 += HALF) {
        uint8_t sfc[HALF], sfn[HALF];
        char2 dv[HALF];
        int32_t sln[HALF];
#pragma unroll
        for (int j = 0; j < HALF; ++j) {
            const int gy = y0 + strip * STRIP + k0 + j;
            sfc[j] = 0;
            dv[j] = make_char2(0, 0);
            if (gx < W && gy < H) {
                const size_t i = (size_t)gy * W + gx;
                sfc[j] = sf[i];
                dv[j] = reinterpret_cast<const char2*>(fd)[i];
            }
        }
#pragma unroll
        for (int j = 0; j < HALF; ++j) {
            const int gy = y0 + strip * STRIP + k0 + j;
            const int nx = gx + dv[j].x, ny = gy + dv[j].y;
            const bool in = gx < W && gy < H && nx >= 0 && ny >= 0 && nx < W && ny < H;
            const size_t ni = in ? (size_t)ny * W + nx : 0;
            sfn[j] = 0;
            sln[j] = 0;
            if (in) {
                sfn[j] = sf[ni];
                sln[j] = slabel[ni];
            }
        }
#pragma unroll
        for (int j = 0; j < HALF; ++j) {
            const int k = k0 + j;
            const int ly = strip * STRIP + k;
            const int gy = y0 + ly;
            uint64_t e = pack(K_STUCK, 0, 0);
            if (gx < W && gy < H) {
                const char2 d = dv[j];
                const int nx = gx + d.x, ny = gy + d.y;
                if (sfc[j] & SF_TRACE) {
                    e = pack(K_RESOLVED, 0, CAGEO_FLOWTRACE_LAKE);
                } else if (nx < 0 || ny < 0 || nx >= W || ny >= H) {
                    e = pack(K_RESOLVED, 1, CAGEO_FLOWTRACE_OUTSIDE);
                } else if (sfn[j] & SF_TRACE) {
                    e = pack(K_RESOLVED, 1, sln[j]);
                } else if (d.x == 0 && d.y == 0) {
                    e = pack(K_STUCK, 0, 0);
                } else if (nx >= x0 && ny >= y0 && nx < x0 + TILE && ny < y0 + TILE) {
                    e = pack(K_LOCAL, 1, (ny - y0) * TILE + (nx - x0));
                    local |= 1u << k;
                } else {
                    e = pack(K_EXIT, 1, (int32_t)((size_t)ny * W + nx));
                }
            }
            L[ly * TILE + lx] = e;
        }
    }
    __syncthreads();

    // Doubling rounds in place: a cell's entry is written by its own thread
    // only, and the target entry it reads is either last round's or already
    // this round's, both exact summaries of the target's path.  So every
    // entry at least doubles its hops a round (paths inside a tile resolve
    // within 13), step counts stay exact sums, and one barrier a round
    // suffices.  Entries are 64-bit LDS words, each read or written whole.
    for (int round = 0; round < 13; ++round) {
        if (!__syncthreads_or(local != 0)) break;
#pragma unroll
        for (int k = 0; k < STRIP; ++k) {
            if (local & (1u << k)) {
                const int li = (strip * STRIP + k) * TILE + lx;
                const uint64_t e = L[li];
                const uint64_t nv = hop(e, L[value_of(e)]);
                L[li] = nv;
                if (kind_of(nv) != K_LOCAL) local &= ~(1u << k);
            }
        }
    }
    __syncthreads();
    // Outputs: the labels of the cells whose path ends in the tile; for an
    // exit cell, the border slot it exits to (in the entry raster) and the
    // mark EXIT_MARK; the tile's border table, every slot written (cells
    // outside the grid and the four unused slots stuck).
    uint64_t* tb = bent + (size_t)t * BORDER_SLOTS;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        const int ly = strip * STRIP + k;
        const int gy = y0 + ly;
        uint64_t e = pack(K_STUCK, 0, 0);
        if (gx < W && gy < H) {
            e = L[ly * TILE + lx];
            const uint32_t kind = kind_of(e);
            const size_t i = (size_t)gy * W + gx;
            int32_t label = 0;
            if (kind == K_RESOLVED) {
                label = value_of(e);
            } else if (kind == K_EXIT) {
                // the first cell outside the tile: a border cell of the next tile
                const uint32_t ti = (uint32_t)value_of(e);
                const int ny = (int)(ti / (uint32_t)W), nx = (int)(ti - (uint32_t)ny * W);
                const int slot = ((ny / TILE) * ntx + nx / TILE) * BORDER_SLOTS +
                                 border_slot(nx % TILE, ny % TILE);
                e = pack(K_EXIT, steps_of(e), slot);
                ent[i] = e;
                label = EXIT_MARK;
            } else {
                e = pack(K_STUCK, 0, 0);  // K_LOCAL: a cycle inside the tile
            }
            lab[i] = label;
        }
        const int slot = border_slot(lx, ly);
        if (slot >= 0) tb[slot] = e;
    }
    if (threadIdx.x < 4)
        tb[threadIdx.x < 2 ? 2 * TILE + (TILE - 1) * threadIdx.x
                           : 3 * TILE + (TILE - 1) * (threadIdx.x - 2)] = pack(K_STUCK, 0, 0);
    __syncthreads();
}


__device__ inline uint32_t ld(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Entries read and rewritten by other waves during the exit rounds move as
// whole 64-bit words.
__device__ inline uint64_t ld64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One raster of a launch.
struct TrRaster {
    const int8_t* fd;
    const uint8_t* sf;
    const int32_t* slabel;
    uint64_t* ent;  // one 64-bit entry per cell (written for exit cells)
    uint64_t* bent; // border table: BORDER_SLOTS entries per tile
    int32_t* lab;
    int W, H, ntx;
    uint32_t* scal;
};

// A launch over one raster (`one`, rs null) or over the nr rasters of a
// table, whose 64x64 blocks are numbered globally (bmap).
struct TrArgs {
    TrRaster one;
    const TrRaster* __restrict__ rs;
    const int2* __restrict__ bmap;  // per block: (raster, block within it)
    int nr, total;
    int max_rounds;
    uint32_t* ctrl;
    GridBar gb;     // grid barrier
};

struct TrBlock {
    TrRaster R;
    int lb;
};
template <bool MULTI>
__device__ inline TrBlock tr_block(const TrArgs& a, int g) {
    TrBlock b;
    g = __builtin_amdgcn_readfirstlane(g);
    if constexpr (!MULTI) {
        b.R = a.one;
        b.lb = g;
    } else {
        const int2 m = a.bmap[g];
        b.R = a.rs[__builtin_amdgcn_readfirstlane(m.x)];
        b.lb = __builtin_amdgcn_readfirstlane(m.y);
    }
    return b;
}

// Final label of cell 0 of a raster (between steps 2 and 3 of the launch).
__device__ inline int32_t label_of_cell0(const TrRaster& R) {
    const int32_t v = R.lab[0];
    if (v != EXIT_MARK) return v;
    const uint64_t f = ld64(&R.bent[value_of(R.ent[0])]);
    return kind_of(f) == K_RESOLVED ? value_of(f) : 0;
}

// Three launches over every raster: the tile pass (one workgroup per
// block), the border-table jumping (persistent, grid barriers between
// rounds) and the exit cells' labels with the connect step (one workgroup
// per block).  The two per-block launches are short-lived and light on
// registers, so many blocks' load chains are in flight at once; only the
// jumping needs co-resident workgroups.

// Tile pass.  Also clears the long-path count: the launch boundary orders
// that before its first use.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_tr_tiles(TrArgs a) {
    __shared__ uint64_t L[TILE * TILE];
    const TrBlock b = tr_block<MULTI>(a, blockIdx.x);
    const TrRaster& R = b.R;
    if (b.lb == 0 && threadIdx.x == 0) R.scal[SC_LONGPATH] = 0;
    trace_tile(L, R.fd, R.sf, R.slabel, R.ent, R.lab, R.bent, R.W, R.H, R.ntx, b.lb);
}

// Jumping over the border tables: thread s takes slot s % BORDER_SLOTS of
// global block s / BORDER_SLOTS (the blocks of one raster are consecutive,
// so a wave's 64 slots lie in one raster).  The barrier ending a round sums
// the workgroups' "an exit remains" bits.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_tr_jump(TrArgs a) {
    if (grid_aborted(a.gb)) return;
    const long long slots = (long long)a.total * BORDER_SLOTS;
    int rounds = 0;
    for (int r = 0; r < a.max_rounds; ++r) {
        rounds = r + 1;
        uint32_t pending = 0;
        for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < slots;
             s += (long long)gridDim.x * blockDim.x) {
            const TrBlock b = tr_block<MULTI>(a, (int)(s / BORDER_SLOTS));
            uint64_t* tb = b.R.bent;
            const size_t own = (size_t)b.lb * BORDER_SLOTS + (size_t)(s % BORDER_SLOTS);
            const uint64_t e = ld64(&tb[own]);
            if (kind_of(e) == K_EXIT) {
                const uint64_t ne = hop(e, ld64(&tb[value_of(e)]));
                pending |= kind_of(ne) == K_EXIT;
                st64(&tb[own], ne);
            }
        }
        const uint32_t mine = __syncthreads_or(pending != 0u) ? 1u : 0u;
        uint32_t any = 0, unused = 0;
        if (!grid_sync_sum(a.gb, 0, mine, 0u, any, unused)) return;
        if (!any) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctrl[CTL_TR_ROUNDS] = (uint32_t)rounds;  // host only
}

// Exit cells take their slot's label; performConnectFlowPath maps a label 0
// to label[0] when that is negative.  An exit whose chain is still moving
// (a cycle through tiles) or ends stuck stays 0.  One workgroup per block,
// a wave per strip of STRIP rows (one lane per column): the strip's labels,
// then its exit cells' entries, then their slots' entries, each round of
// loads issued whole before the next.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_tr_exits(TrArgs a) {
    const TrBlock b = tr_block<MULTI>(a, blockIdx.x);
    const TrRaster& R = b.R;
    const int x = (b.lb % R.ntx) * TILE + (threadIdx.x & 63);
    const int y0 = (b.lb / R.ntx) * TILE + (int)(threadIdx.x >> 6) * STRIP;
    const int rows = x < R.W ? min(STRIP, R.H - y0) : 0;
    const int32_t l0 = label_of_cell0(R);
    int32_t v[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k) v[k] = k < rows ? R.lab[(size_t)(y0 + k) * R.W + x] : 1;
    uint64_t e[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k)
        e[k] = v[k] == EXIT_MARK ? R.ent[(size_t)(y0 + k) * R.W + x] : 0;
    uint64_t f[STRIP];
#pragma unroll
    for (int k = 0; k < STRIP; ++k) f[k] = v[k] == EXIT_MARK ? ld64(&R.bent[value_of(e[k])]) : 0;
    uint32_t longp = 0;
#pragma unroll
    for (int k = 0; k < STRIP; ++k) {
        if (k >= rows) continue;
        int32_t w = v[k];
        if (w == EXIT_MARK) {
            w = 0;
            if (kind_of(f[k]) == K_RESOLVED) {
                w = value_of(f[k]);
                if (steps_of(e[k]) + steps_of(f[k]) > 100001u) ++longp;  // past the walk guard
            }
        }
        if (w == 0 && l0 < 0) w = l0;
        if (w != v[k]) R.lab[(size_t)(y0 + k) * R.W + x] = w;
    }
    if (__any(longp)) {
        for (int off = 32; off > 0; off >>= 1) longp += (uint32_t)__shfl_xor((int)longp, off);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&R.scal[SC_LONGPATH], longp);
            atomicAdd(&R.scal[SC_LONGRUN], longp);
        }
    }
}

// The jumping's co-resident workgroups: at most two per CU; with a few
// thousand slots per round more workgroups would only lengthen its
// barriers.
int jump_cap(const void* fn, int device, int* cap) {
    int per_cu = 0;
    CAGEO_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TILE_THREADS, 0));
    hipDeviceProp_t prop;
    CAGEO_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (per_cu < 1) return CAGEO_ERR_HIP;
    *cap = std::min(per_cu, 2) * prop.multiProcessorCount;
    return CAGEO_OK;
}

int max_rounds_for(size_t cells) {
    int m = 2;
    for (size_t k = 1; k < cells; k <<= 1) ++m;
    return m;
}

TrRaster tr_raster(Context* c, const int8_t* fd, const uint8_t* sflags, const int32_t* slabel,
                   int32_t* lab) {
    TrRaster R;
    R.fd = fd;
    R.sf = sflags;
    R.slabel = slabel;
    R.ent = c->d_ptr;
    R.bent = c->d_bent;
    R.lab = lab;
    R.W = c->W;
    R.H = c->H;
    R.ntx = c->ntx;
    R.scal = c->d_scal;
    return R;
}

}  // namespace

// Needs one 64-bit raster, c->d_ptr.  The long-path count is
// left in the scalar block.
// Co-resident workgroups of the persistent kernel on this device.
int flow_tracing_grid_cap(Context* c, int* cap) {
    return jump_cap(reinterpret_cast<const void*>(k_tr_jump<false>), c->device, cap);
}

int flow_tracing(Context* c, const int8_t* fd, const uint8_t* sflags, const int32_t* slabel,
                 int32_t* lab, int64_t* long_paths) {
    if (c->tr_grid <= 0) CAGEO_TRY(flow_tracing_grid_cap(c, &c->tr_grid));
    TrArgs a;
    a.one = tr_raster(c, fd, sflags, slabel, lab);
    a.rs = nullptr;
    a.bmap = nullptr;
    a.nr = 1;
    a.total = c->ntiles;
    a.max_rounds = max_rounds_for(c->n);
    a.ctrl = c->d_ctrl;
    a.gb = context_gridbar(c);
    hipLaunchKernelGGL(k_tr_tiles<false>, dim3(a.total), dim3(TILE_THREADS), 0, c->stream, a);
    // ordinary launch within the occupancy cap (see pit_fill_quantized)
    hipLaunchKernelGGL(k_tr_jump<false>, dim3(c->tr_grid), dim3(TILE_THREADS), 0, c->stream, a);
    hipLaunchKernelGGL(k_tr_exits<false>, dim3(a.total), dim3(TILE_THREADS), 0, c->stream, a);
    CAGEO_HIP_TRY(hipGetLastError());
    if (long_paths) {
        CAGEO_TRY(read_scalars(c));
        *long_paths = (int64_t)c->h_scal[SC_LONGPATH];
    }
    return CAGEO_OK;
}

// Raster table of a batch: each context's d_fd -> d_lab with its stream
// flags and labels.
int flow_tracing_tables(MultiState* m, Context* const* ctxs) {
    std::vector<TrRaster> t(m->nr);
    for (int r = 0; r < m->nr; ++r) {
        Context* c = ctxs[r];
        t[r] = tr_raster(c, c->d_fd, c->d_sflags, c->d_slabel, c->d_lab);
    }
    CAGEO_HIP_TRY(hipMalloc(&m->d_tr, t.size() * sizeof(TrRaster)));
    CAGEO_HIP_TRY(hipMemcpy(m->d_tr, t.data(), t.size() * sizeof(TrRaster), hipMemcpyHostToDevice));
    CAGEO_TRY(jump_cap(reinterpret_cast<const void*>(k_tr_jump<true>), ctxs[0]->device, &m->tr_cap));
    if (m->tr_grid <= 0) m->tr_grid = m->tr_cap;
    return CAGEO_OK;
}

int flow_tracing_multi(MultiState* m, hipStream_t s) {
    TrArgs a;
    a.one = TrRaster{};
    a.rs = static_cast<const TrRaster*>(m->d_tr);
    a.bmap = m->d_bmap;
    a.nr = m->nr;
    a.total = m->total;
    a.max_rounds = max_rounds_for((size_t)m->max_cells);
    a.ctrl = m->d_ctrl;
    a.gb = multi_gridbar(m);
    hipLaunchKernelGGL(k_tr_tiles<true>, dim3(a.total), dim3(TILE_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_tr_jump<true>, dim3(m->tr_grid), dim3(TILE_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_tr_exits<true>, dim3(a.total), dim3(TILE_THREADS), 0, s, a);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
