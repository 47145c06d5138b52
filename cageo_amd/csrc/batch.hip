// Batched Monte Carlo: many DEM tiles of uneven size x many realizations.
//
// No reference counterpart: the reference runs one DEM per process
// (main.cpp:180-283).  Each tile gets its own context (workspace sized to the
// tile, all resident in HBM), its own error model and filter, and its own
// Philox key.  Tile t, realization r is the same computation as
// cageo_run_realizations on that tile alone, so every mode below gives the
// same counter rasters:
//   PER_TILE  one blocking run_realizations call per tile (the baseline);
//   STREAM    per realization: the random field, convolution and burning of
//             every tile, then ONE pit-fill, ONE flow-routing and ONE
//             flow-tracing launch over the 64x64 blocks of all tiles at once
//             (the persistent kernels take a table of rasters), then the
//             border accumulation of every tile; all on one stream, one wait;
//   GRAPH     that sequence for one realization captured once as a hipGraph
//             and replayed n times; the random-field kernels read the
//             realization index from device memory, which a last node bumps.
// A small tile has few 64x64 blocks, so its own persistent launch leaves most
// of the chip idle while it iterates; in one launch over all tiles the number
// of barrier-separated passes is that of the slowest tile, not the sum.
// The autotuner picks, per tile, the workgroup count of its own persistent
// launches and the row-block height of the 2-D convolution, and then the
// workgroup count of the all-tile launches.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace cageo {

namespace {

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
__global__ void k_bump_u32(uint32_t* p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += 1u;
}

}  // namespace

struct BatchTile {
    cageo_context* ctx = nullptr;
    uint64_t seed = 0;
    bool ready = false;
};

}  // namespace cageo

struct cageo_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;  // created here: the NULL stream cannot be captured
    std::vector<cageo::BatchTile> tiles;
    uint32_t* d_iter = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<float*> graph_mc;  // counter rasters baked into the graph
    cageo::MultiState multi;       // all-tile persistent launches
    bool multi_ready = false;
    int multi_grid = 0;            // tuned workgroups of the all-tile launches (0: cap)
};

namespace cageo {

namespace {

void drop_graph(cageo_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->graph_mc.clear();
}

// One realization of tile t: random field (realization = first + base) and
// the per-realization stages, accumulated into mc.
int enqueue_tile(cageo_batch* b, int t, float* mc, uint32_t first, const uint32_t* base) {
    Context* c = &b->tiles[t].ctx->c;
    CAGEO_TRY(launch_random_field(c, c->d_a, b->tiles[t].seed, first, 0u, c->mean, c->std_dev,
                                  base));
    return realization_body(c, mc, nullptr);
}

void release_multi(cageo_batch* b) {
    if (b->multi_ready) multi_release(&b->multi);
    b->multi_ready = false;
}

// Device state and raster tables of the all-tile launches (once the tiles
// are set up; dropped when one is set up again).
int ensure_multi(cageo_batch* b) {
    if (b->multi_ready) return CAGEO_OK;
    std::vector<Context*> ctxs;
    for (auto& t : b->tiles) ctxs.push_back(&t.ctx->c);
    CAGEO_TRY(multi_setup(&b->multi, ctxs.data(), (int)ctxs.size()));
    b->multi_ready = true;
    if (b->multi_grid > 0)
        b->multi.pf_grid = b->multi.fr_grid = b->multi.tr_grid = b->multi_grid;
    return CAGEO_OK;
}

// One realization of every tile with the all-tile persistent launches.
int enqueue_all(cageo_batch* b, float* const* mc, uint32_t first, const uint32_t* base) {
    const int T = (int)b->tiles.size();
    for (int t = 0; t < T; ++t) {
        Context* c = &b->tiles[t].ctx->c;
        CAGEO_TRY(launch_random_field(c, c->d_a, b->tiles[t].seed, first, 0u, c->mean,
                                      c->std_dev, base));
        CAGEO_TRY(realization_pre(c, nullptr));
    }
    CAGEO_TRY(pit_fill_multi(&b->multi, b->stream));
    CAGEO_TRY(flow_routing_multi(&b->multi, b->stream));
    CAGEO_TRY(flow_tracing_multi(&b->multi, b->stream));
    for (int t = 0; t < T; ++t)
        if (mc[t]) CAGEO_TRY(launch_border(&b->tiles[t].ctx->c, mc[t], b->tiles[t].ctx->c.d_lab));
    return CAGEO_OK;
}

int check_all(cageo_batch* b) {
    CAGEO_HIP_TRY(hipStreamSynchronize(b->stream));
    for (auto& t : b->tiles) CAGEO_TRY(check_status(&t.ctx->c));
    if (b->multi_ready) {
        uint32_t st = 0;
        CAGEO_HIP_TRY(hipMemcpy(&st, b->multi.d_ctrl + CTL_STATUS, sizeof(st),
                                hipMemcpyDeviceToHost));
        if (st) return -(int)st;
    }
    return CAGEO_OK;
}

int clear_all(cageo_batch* b) {
    for (auto& t : b->tiles) CAGEO_TRY(clear_status(&t.ctx->c));
    if (b->multi_ready)
        CAGEO_TRY(reset_launch_state(b->stream, b->multi.d_ctrl + CTL_STATUS, b->multi.d_bar));
    return CAGEO_OK;
}

int build_graph(cageo_batch* b, float* const* mc) {
    drop_graph(b);
    // tables and workgroup counts are fixed before capture (no allocations or
    // device queries inside it)
    CAGEO_TRY(ensure_multi(b));
    hipGraph_t g = nullptr;
    CAGEO_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    int st = enqueue_all(b, mc, 0u, b->d_iter);
    if (st == CAGEO_OK) {
        hipLaunchKernelGGL(k_bump_u32, dim3(1), dim3(64), 0, b->stream, b->d_iter);
        if (hipGetLastError() != hipSuccess) st = CAGEO_ERR_HIP;
    }
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != CAGEO_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st != CAGEO_OK ? st : CAGEO_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        b->exec = nullptr;
        return CAGEO_ERR_HIP;
    }
    b->graph_mc.assign(mc, mc + b->tiles.size());
    return CAGEO_OK;
}

// Device time of `reps` realizations of tile t with the current settings.
int time_tile(cageo_batch* b, int t, float* mc, int reps, float* ms) {
    Context* c = &b->tiles[t].ctx->c;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    CAGEO_HIP_TRY(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; ++r) CAGEO_TRY(enqueue_tile(b, t, mc, (uint32_t)r, nullptr));
    CAGEO_HIP_TRY(hipEventRecord(e1, b->stream));
    CAGEO_HIP_TRY(hipEventSynchronize(e1));
    CAGEO_TRY(check_status(c));
    CAGEO_HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    return CAGEO_OK;
}

// Device time of `reps` realizations of every tile with the all-tile
// launches at the current workgroup count.
int time_all(cageo_batch* b, float* const* mc, int reps, float* ms) {
    Context* c = &b->tiles[0].ctx->c;
    hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
    CAGEO_HIP_TRY(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; ++r) CAGEO_TRY(enqueue_all(b, mc, (uint32_t)r, nullptr));
    CAGEO_HIP_TRY(hipEventRecord(e1, b->stream));
    CAGEO_HIP_TRY(hipEventSynchronize(e1));
    CAGEO_TRY(check_all(b));
    CAGEO_HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    return CAGEO_OK;
}

}  // namespace

template <class T>
static int dev_alloc(T** p, size_t count) {
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) + 16) != hipSuccess) {
        *p = nullptr;
        return CAGEO_ERR_NOMEM;
    }
    return CAGEO_OK;
}

int multi_setup(MultiState* m, Context* const* ctxs, int nr) {
    *m = MultiState{};
    m->nr = nr;
    std::vector<int> base(nr + 1, 0);
    for (int r = 0; r < nr; ++r) {
        base[r + 1] = base[r] + ctxs[r]->ntiles;
        m->max_cells = std::max(m->max_cells, (int)ctxs[r]->n);
    }
    m->total = base[nr];
    int st = CAGEO_OK;
    auto A = [&](int s) {
        if (st == CAGEO_OK) st = s;
    };
    A(dev_alloc(&m->d_base, (size_t)nr + 1));
    A(dev_alloc(&m->d_bmap, (size_t)m->total));
    A(dev_alloc(&m->d_rinfo, (size_t)nr));
    A(dev_alloc(&m->d_work, 7 * (size_t)m->total));
    A(dev_alloc(&m->d_mark_pf, (size_t)m->total));
    A(dev_alloc(&m->d_mark_fr, (size_t)m->total));
    A(dev_alloc(&m->d_rows, 4 * (size_t)m->total));
    A(dev_alloc(&m->d_topen, (size_t)m->total));
    A(dev_alloc(&m->d_tcost, (size_t)m->total));
    A(dev_alloc(&m->d_ctrl, (size_t)CTL_WORDS));
    A(dev_alloc(&m->d_cert, 3 * (size_t)nr));
    A(dev_alloc(&m->d_bar, (size_t)BAR_WORDS));
    std::vector<int2> bmap((size_t)m->total);
    for (int r = 0; r < nr; ++r)
        for (int k = base[r]; k < base[r + 1]; ++k) bmap[k] = make_int2(r, k - base[r]);
    if (st == CAGEO_OK &&
        (hipMemcpy(m->d_base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice) !=
             hipSuccess ||
         hipMemcpy(m->d_bmap, bmap.data(), bmap.size() * sizeof(int2), hipMemcpyHostToDevice) !=
             hipSuccess ||
         hipMemset(m->d_mark_pf, 0, (size_t)m->total * sizeof(uint32_t)) != hipSuccess ||
         hipMemset(m->d_mark_fr, 0, (size_t)m->total * sizeof(uint32_t)) != hipSuccess ||
         hipMemset(m->d_ctrl, 0, CTL_WORDS * sizeof(uint32_t)) != hipSuccess ||
         hipMemset(m->d_cert, 0, 3 * (size_t)nr * sizeof(uint32_t)) != hipSuccess ||
         hipMemset(m->d_bar, 0, BAR_WORDS * sizeof(uint32_t)) != hipSuccess))
        st = CAGEO_ERR_HIP;
    // the all-tile launches share the abort word and wait bound of tile 0's
    // context (cageo_batch_abort raises every tile's word)
    m->d_habort = ctxs[0]->d_habort;
    m->bar_limit = ctxs[0]->bar_limit;
    if (st == CAGEO_OK) st = pit_fill_tables(m, ctxs);
    if (st == CAGEO_OK) st = flow_routing_tables(m, ctxs);
    if (st == CAGEO_OK) st = flow_tracing_tables(m, ctxs);
    if (st != CAGEO_OK) multi_release(m);
    return st;
}

void multi_release(MultiState* m) {
    void* ptrs[] = {m->d_base, m->d_bmap, m->d_rinfo, m->d_work, m->d_mark_pf, m->d_mark_fr, m->d_rows, m->d_topen, m->d_tcost, m->d_ctrl,
                    m->d_cert, m->d_bar, m->d_pf, m->d_fr, m->d_tr};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    *m = MultiState{};
}

}  // namespace cageo

using cageo::Context;

extern "C" {

int cageo_batch_create(cageo_batch** out, int ntiles, const int* widths, const int* heights,
                       void* stream) {
    if (!out || ntiles <= 0 || !widths || !heights) return CAGEO_ERR_ARG;
    cageo_batch* b = new cageo_batch();
    b->stream = reinterpret_cast<hipStream_t>(stream);
    if (!b->stream) {
        // a blocking stream (default flags and priority): it still orders
        // after work on the NULL stream, and unlike it, it can be captured
        if (hipStreamCreate(&b->stream) != hipSuccess) {
            delete b;
            return CAGEO_ERR_HIP;
        }
        b->own_stream = true;
    }
    b->tiles.resize(ntiles);
    int st = CAGEO_OK;
    for (int t = 0; t < ntiles && st == CAGEO_OK; ++t)
        st = cageo_context_create(&b->tiles[t].ctx, widths[t], heights[t], b->stream);
    if (st == CAGEO_OK &&
        hipMalloc(reinterpret_cast<void**>(&b->d_iter), sizeof(uint32_t)) != hipSuccess)
        st = CAGEO_ERR_NOMEM;
    if (st != CAGEO_OK) {
        for (auto& t : b->tiles)
            if (t.ctx) cageo_context_destroy(t.ctx);
        if (b->own_stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return st;
    }
    *out = b;
    return CAGEO_OK;
}

int cageo_batch_destroy(cageo_batch* b) {
    if (!b) return CAGEO_ERR_ARG;
    (void)hipStreamSynchronize(b->stream);
    cageo::drop_graph(b);
    cageo::release_multi(b);
    for (auto& t : b->tiles)
        if (t.ctx) cageo_context_destroy(t.ctx);
    if (b->d_iter) (void)hipFree(b->d_iter);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CAGEO_OK;
}

int cageo_batch_context(cageo_batch* b, int tile, cageo_context** ctx) {
    if (!b || !ctx || tile < 0 || tile >= (int)b->tiles.size()) return CAGEO_ERR_ARG;
    *ctx = b->tiles[tile].ctx;
    return CAGEO_OK;
}

int cageo_batch_setup_tile(cageo_batch* b, int tile, const float* d_dem, const float* d_stream,
                           float nodata_dem, float nodata_stream, int bevel_radius, int model,
                           float practical_range, float cell_size, float mean, float std_dev,
                           uint64_t seed) {
    if (!b || tile < 0 || tile >= (int)b->tiles.size()) return CAGEO_ERR_ARG;
    cageo_context* ctx = b->tiles[tile].ctx;
    const int r = cageo_filter_radius(practical_range, cell_size);
    if (r < 0) return CAGEO_ERR_ARG;
    const int L = 2 * r + 1;
    std::vector<float> taps(model == CAGEO_FILTER_GAUSSIAN ? (size_t)L : (size_t)L * L);
    CAGEO_TRY(cageo_create_filter(model, practical_range, cell_size, r, taps.data()));
    CAGEO_TRY(cageo_set_filter(ctx, model, r, taps.data()));
    CAGEO_TRY(cageo_mc_setup(ctx, d_dem, d_stream, nodata_dem, nodata_stream, bevel_radius, mean,
                             std_dev));
    b->tiles[tile].seed = seed;
    b->tiles[tile].ready = true;
    cageo::drop_graph(b);
    cageo::release_multi(b);  // the table holds the tile's nodata value
    return CAGEO_OK;
}

int cageo_batch_run(cageo_batch* b, float* const* d_mc, int n, uint32_t first, int mode) {
    if (!b || !d_mc || n < 0) return CAGEO_ERR_ARG;
    for (auto& t : b->tiles)
        if (!t.ready) return CAGEO_ERR_STATE;
    if (mode != CAGEO_BATCH_PER_TILE) CAGEO_TRY(cageo::ensure_multi(b));
    CAGEO_TRY(cageo::clear_all(b));
    const int T = (int)b->tiles.size();
    if (mode == CAGEO_BATCH_PER_TILE) {
        for (int t = 0; t < T; ++t)
            CAGEO_TRY(cageo_run_realizations(b->tiles[t].ctx, d_mc[t], n, b->tiles[t].seed, first));
        return CAGEO_OK;
    }
    if (mode == CAGEO_BATCH_STREAM) {
        for (int r = 0; r < n; ++r) CAGEO_TRY(cageo::enqueue_all(b, d_mc, first + (uint32_t)r, nullptr));
        return cageo::check_all(b);
    }
    if (mode != CAGEO_BATCH_GRAPH) return CAGEO_ERR_ARG;
    bool same = b->exec && b->graph_mc.size() == (size_t)T;
    for (int t = 0; same && t < T; ++t) same = b->graph_mc[t] == d_mc[t];
    if (!same) CAGEO_TRY(cageo::build_graph(b, d_mc));
    hipLaunchKernelGGL(cageo::k_set_u32, dim3(1), dim3(64), 0, b->stream, b->d_iter, first);
    CAGEO_HIP_TRY(hipGetLastError());
    for (int r = 0; r < n; ++r) CAGEO_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    return cageo::check_all(b);
}

int cageo_batch_tune(cageo_batch* b, float* const* d_scratch_mc, int reps, int* out_config) {
    if (!b || !d_scratch_mc || reps < 1) return CAGEO_ERR_ARG;
    for (auto& t : b->tiles)
        if (!t.ready) return CAGEO_ERR_STATE;
    cageo::drop_graph(b);
    for (int t = 0; t < (int)b->tiles.size(); ++t) {
        Context* c = &b->tiles[t].ctx->c;
        int caps[3];
        CAGEO_TRY(cageo::pit_fill_grid_cap(c, &caps[0]));
        CAGEO_TRY(cageo::flow_routing_grid_cap(c, &caps[1]));
        CAGEO_TRY(cageo::flow_tracing_grid_cap(c, &caps[2]));
        const int cap = std::min(caps[0], std::min(caps[1], caps[2]));
        // candidate workgroup counts: around the tile count, and fractions of
        // the full chip
        std::vector<int> cand;
        for (int g : {c->ntiles / 2, c->ntiles, 2 * c->ntiles, cap / 8, cap / 4, cap / 2, cap})
            if (g >= 1 && g <= cap && std::find(cand.begin(), cand.end(), g) == cand.end())
                cand.push_back(g);
        float best = 3.4e38f;
        int best_g = cap;
        for (int g : cand) {
            c->pf_grid = c->fr_grid = c->tr_grid = g;
            float ms = 0.0f;
            CAGEO_TRY(cageo::time_tile(b, t, d_scratch_mc[t], 1, &ms));  // warm
            CAGEO_TRY(cageo::time_tile(b, t, d_scratch_mc[t], reps, &ms));
            if (ms < best) {
                best = ms;
                best_g = g;
            }
        }
        c->pf_grid = c->fr_grid = c->tr_grid = best_g;
        int best_ty = 0;
        if (c->model == CAGEO_FILTER_EXPONENTIAL) {
            float bt = 3.4e38f;
            for (int ty : {16, 8, 4}) {
                int ty0, S;
                size_t lds;
                CAGEO_TRY(cageo::conv_2d_config(c->radius, &ty0, &S, &lds));
                if ((size_t)(4 * ty + 2 * c->radius) * S * sizeof(float) > 160 * 1024) continue;
                c->conv_ty = ty;
                float ms = 0.0f;
                CAGEO_TRY(cageo::time_tile(b, t, d_scratch_mc[t], reps, &ms));
                if (ms < bt) {
                    bt = ms;
                    best_ty = ty;
                }
            }
        }
        c->conv_ty = best_ty;
        if (out_config) {
            out_config[2 * t] = best_g;
            out_config[2 * t + 1] = best_ty;
        }
    }
    // workgroups of the all-tile launches: fractions of the smallest
    // co-resident count of the all-tile kernels themselves
    CAGEO_TRY(cageo::ensure_multi(b));
    cageo::MultiState& m = b->multi;
    const int cap = std::min(m.pf_cap, std::min(m.fr_cap, m.tr_cap));
    float best = 3.4e38f;
    int best_g = cap;
    for (int g : {cap / 4, cap / 2, (3 * cap) / 4, cap}) {
        if (g < 1) continue;
        m.pf_grid = m.fr_grid = m.tr_grid = g;
        float ms = 0.0f;
        CAGEO_TRY(cageo::time_all(b, d_scratch_mc, 1, &ms));  // warm
        CAGEO_TRY(cageo::time_all(b, d_scratch_mc, reps, &ms));
        if (ms < best) {
            best = ms;
            best_g = g;
        }
    }
    b->multi_grid = best_g;
    m.pf_grid = m.fr_grid = m.tr_grid = best_g;
    return CAGEO_OK;
}

int cageo_batch_info(cageo_batch* b, int* out) {
    if (!b || !out) return CAGEO_ERR_ARG;
    for (auto& t : b->tiles)
        if (!t.ready) return CAGEO_ERR_STATE;
    CAGEO_TRY(cageo::ensure_multi(b));
    out[0] = b->multi.total;
    out[1] = b->multi.pf_grid;
    out[2] = b->multi.fr_grid;
    out[3] = b->multi.tr_grid;
    return CAGEO_OK;
}

int cageo_batch_abort(cageo_batch* b, int set) {
    if (!b) return CAGEO_ERR_ARG;
    for (auto& t : b->tiles) CAGEO_TRY(cageo_context_abort(t.ctx, set));
    return CAGEO_OK;
}

int cageo_batch_configure_tile(cageo_batch* b, int tile, int persistent_workgroups, int conv_rows) {
    if (!b || tile < 0 || tile >= (int)b->tiles.size()) return CAGEO_ERR_ARG;
    CAGEO_TRY(cageo_context_configure(b->tiles[tile].ctx, persistent_workgroups, conv_rows));
    cageo::drop_graph(b);
    return CAGEO_OK;
}

}  // extern "C"
