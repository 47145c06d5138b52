// C ABI of libcageo_hip: context (device workspace), stage entry points that
// replace the reference's XAlgorithm::execute calls, and the fused Monte Carlo
// driver that replaces the loop body of main.cpp:180-283.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace cageo {

// min <- max ordered value, max and "bad" <- 0.  A kernel rather than memsets
// so that a captured graph holds no memset nodes.
__global__ void k_reset_minmax(uint32_t* __restrict__ scal) {
    if (threadIdx.x == 0) {
        scal[SC_MIN] = 0xffffffffu;
        scal[SC_MAX] = 0u;
        scal[SC_BAD] = 0u;
    }
}

int reset_minmax(Context* c) {
    hipLaunchKernelGGL(k_reset_minmax, dim3(1), dim3(64), 0, c->stream, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

namespace {

int sync(Context* c) {
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    return CAGEO_OK;
}

template <class T>
int alloc(T** p, size_t count) {
    if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)) !=
        hipSuccess) {
        *p = nullptr;
        return CAGEO_ERR_NOMEM;
    }
    return CAGEO_OK;
}

void free_all(Context* c) {
    void* ptrs[] = {c->d_mark, c->d_mark_pf, c->d_topen, c->d_tcost, c->d_rows, c->d_work3, c->d_ctrl, c->d_bar, c->d_scal, c->d_bmm, c->d_filter,
                    c->d_sflags, c->d_slabel, c->d_off, c->d_factor, c->d_a, c->d_b, c->d_z,
                    c->d_w, c->d_st, c->d_fd, c->d_dist, c->d_lab, c->d_ptr, c->d_bent,
                    c->d_lr_col, c->d_lr_row, c->d_lr_planes};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    release_fft(c);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->h_abort) (void)hipHostFree(c->h_abort);
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
    c->evpool.clear();
}

}  // namespace

// Convolution, realistic DEM, burning and quantisation of the error field
// in c->d_a into c->d_z (with evs: events evs[2], evs[3] after the first two).
int realization_pre(Context* c, hipEvent_t* evs) {
    float* e = nullptr;
    CAGEO_TRY(launch_convolve_to(c, c->d_a, c->d_b, &e));
    if (evs) CAGEO_HIP_TRY(hipEventRecord(evs[2], c->stream));
    CAGEO_TRY(launch_realistic_burn_quant(c, c->dem, e, c->d_z));  // min/max words included
    if (evs) CAGEO_HIP_TRY(hipEventRecord(evs[3], c->stream));
    return CAGEO_OK;
}

// One realization from the error field already in c->d_a (stages a3..a9).
// With `evs` (8 events), evs[2..7] are recorded after each stage.
int realization_body(Context* c, float* mc, hipEvent_t* evs) {
    int ev = 2;
    auto mark = [&]() -> int {
        if (evs) CAGEO_HIP_TRY(hipEventRecord(evs[ev++], c->stream));
        return CAGEO_OK;
    };
    CAGEO_TRY(realization_pre(c, evs));
    ev = 4;
    CAGEO_TRY(pit_fill_quantized(c, c->d_z, c->d_w, c->nodata_dem, true));
    CAGEO_TRY(mark());
    CAGEO_TRY(flow_routing(c, c->d_w, c->d_fd));
    CAGEO_TRY(mark());
    CAGEO_TRY(flow_tracing(c, c->d_fd, c->d_sflags, c->d_slabel, c->d_lab, nullptr));
    CAGEO_TRY(mark());
    if (mc) CAGEO_TRY(launch_border(c, mc, c->d_lab));
    CAGEO_TRY(mark());
    return CAGEO_OK;
}


int read_scalars(Context* c) {
    CAGEO_HIP_TRY(hipMemcpyAsync(c->h_scal, c->d_scal, SC_NSLOTS * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    return CAGEO_OK;
}

__global__ void k_reset_launch_state(uint32_t* __restrict__ status, uint32_t* __restrict__ bar,
                                     uint32_t* __restrict__ scal) {
    // the generation word (64-bit) keeps counting; every other barrier word restarts
    for (int i = threadIdx.x; i < BAR_WORDS; i += blockDim.x)
        if (i != BAR_GEN && i != BAR_GEN + 1) bar[i] = 0u;
    if (threadIdx.x == 0) {
        *status = 0u;
        if (scal) scal[SC_LONGRUN] = 0u;
    }
}

int reset_launch_state(hipStream_t s, uint32_t* status, uint32_t* bar, uint32_t* scal) {
    hipLaunchKernelGGL(k_reset_launch_state, dim3(1), dim3(256), 0, s, status, bar, scal);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int barrier_ticks(int device, double seconds, uint64_t* ticks) {
    int khz = 0;
    CAGEO_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    if (khz <= 0) khz = 100000;  // the 100 MHz constant clock of CDNA parts
    const double t = seconds * 1e3 * (double)khz;
    *ticks = t >= 1.8e19 ? ~0ull : (uint64_t)t;
    return CAGEO_OK;
}

int clear_status(Context* c) {
    return reset_launch_state(c->stream, c->d_ctrl + CTL_STATUS, c->d_bar, c->d_scal);
}

// Synchronises and returns the sticky device status of the persistent kernels.
int check_status(Context* c) {
    uint32_t st = 0;
    CAGEO_HIP_TRY(hipMemcpyAsync(&c->h_scal[SC_NSLOTS - 1], c->d_ctrl + CTL_STATUS,
                                 sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    st = c->h_scal[SC_NSLOTS - 1];
    return st ? -(int)st : CAGEO_OK;
}

// Low-rank separable form of the exponential filter for the fast mode:
// terms until the remainder is below 1e-6 of the filter (at most 32),
// column taps padded to whole register-ring turns, row taps lambda_k v_k.
int prepare_low_rank(Context* c) {
    if (c->model != CAGEO_FILTER_EXPONENTIAL || c->h_taps.empty()) return CAGEO_OK;
    // the FFT form where the grid suits it; the low-rank form otherwise
    const int fst = prepare_fft(c);
    if (fst == CAGEO_OK) {
        if (c->d_lr_col) (void)hipFree(c->d_lr_col);
        if (c->d_lr_row) (void)hipFree(c->d_lr_row);
        if (c->d_lr_planes) (void)hipFree(c->d_lr_planes);
        c->d_lr_col = c->d_lr_row = c->d_lr_planes = nullptr;
        c->lr_k = 0;
        return CAGEO_OK;
    }
    if (fst != CAGEO_ERR_RANGE) return fst;
    const int r = c->radius, L = 2 * r + 1, KMAX = 32;
    int LC, LP, S;
    size_t lc, lr;
    if (lr_cols_geometry(r, &LC, &lc) != CAGEO_OK || lr_rows_geometry(r, &LP, &S, &lr) != CAGEO_OK)
        return CAGEO_ERR_ARG;
    std::vector<float> vec((size_t)KMAX * L), val(KMAX);
    int K = 0;
    double resid = 0.0;
    CAGEO_TRY(cageo_filter_low_rank(c->h_taps.data(), r, KMAX, 1e-6, vec.data(), val.data(), &K,
                                    &resid));
    // row taps carry 4 zero floats after the last term: the row pass reads
    // one tap quadruple ahead
    std::vector<float> col((size_t)K * LC, 0.0f), row((size_t)K * LP + 4, 0.0f);
    for (int k = 0; k < K; ++k)
        for (int i = 0; i < L; ++i) {
            col[(size_t)k * LC + i] = vec[(size_t)k * L + i];
            row[(size_t)k * LP + i] = val[k] * vec[(size_t)k * L + i];
        }
    if (c->d_lr_col) (void)hipFree(c->d_lr_col);
    if (c->d_lr_row) (void)hipFree(c->d_lr_row);
    if (c->d_lr_planes) (void)hipFree(c->d_lr_planes);
    c->d_lr_col = c->d_lr_row = c->d_lr_planes = nullptr;
    c->lr_k = 0;
    CAGEO_TRY(alloc(&c->d_lr_col, col.size()));
    CAGEO_TRY(alloc(&c->d_lr_row, row.size()));
    CAGEO_TRY(alloc(&c->d_lr_planes, (size_t)K * c->n));
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_lr_col, col.data(), col.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_lr_row, row.data(), row.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    c->lr_k = K;
    return CAGEO_OK;
}


}  // namespace cageo

using cageo::Context;

extern "C" {

const char* cageo_status_string(int s) {
    switch (s) {
        case CAGEO_OK: return "ok";
        case CAGEO_ERR_ARG: return "invalid argument";
        case CAGEO_ERR_HIP: return "HIP runtime error";
        case CAGEO_ERR_NOMEM: return "device allocation failed";
        case CAGEO_ERR_NOT_CONVERGED: return "iteration cap reached";
        case CAGEO_ERR_STATE: return "invalid call order";
        case CAGEO_ERR_RANGE: return "input outside the supported domain";
        case CAGEO_ERR_ABORTED: return "persistent launch abandoned (abort word or barrier time limit)";
        default: return "unknown status";
    }
}

const char* cageo_version(void) { return "cageo_hip 0.1 gfx950"; }

int cageo_context_create(cageo_context** out, int W, int H, void* stream) {
    if (!out || W <= 0 || H <= 0 || (int64_t)W * H >= (int64_t)1 << 31) return CAGEO_ERR_ARG;
    cageo_context* ctx = new cageo_context();
    Context* c = &ctx->c;
    c->W = W;
    c->H = H;
    c->n = (size_t)W * H;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete ctx;
        return CAGEO_ERR_HIP;
    }
    c->ntx = (W + cageo::TILE - 1) / cageo::TILE;
    c->nty = (H + cageo::TILE - 1) / cageo::TILE;
    c->ntiles = c->ntx * c->nty;
    const size_t n = c->n;
    int st = CAGEO_OK;
    auto A = [&](int s) {
        if (st == CAGEO_OK) st = s;
    };
    A(cageo::alloc(&c->d_mark, c->ntiles));
    A(cageo::alloc(&c->d_mark_pf, c->ntiles));
    A(cageo::alloc(&c->d_rows, 4 * (size_t)c->ntiles));
    A(cageo::alloc(&c->d_topen, c->ntiles));
    A(cageo::alloc(&c->d_tcost, c->ntiles));
    A(cageo::alloc(&c->d_work3, 7 * (size_t)c->ntiles));
    A(cageo::alloc(&c->d_ctrl, cageo::CTL_WORDS));
    A(cageo::alloc(&c->d_bar, cageo::BAR_WORDS));
    A(cageo::alloc(&c->d_scal, cageo::SC_NSLOTS));
    A(cageo::alloc(&c->d_bmm, 3 * cageo::burn_blocks(W, H)));
    A(cageo::alloc(&c->d_sflags, n));
    A(cageo::alloc(&c->d_slabel, n));
    const int maxoff = (2 * CAGEO_MAX_BEVEL_RADIUS + 1) * (2 * CAGEO_MAX_BEVEL_RADIUS + 1);
    A(cageo::alloc(&c->d_off, 2 * maxoff));
    A(cageo::alloc(&c->d_factor, maxoff));
    A(cageo::alloc(&c->d_a, n));
    A(cageo::alloc(&c->d_b, n));
    A(cageo::alloc(&c->d_z, n));
    A(cageo::alloc(&c->d_w, n));
    A(cageo::alloc(&c->d_st, n));
    A(cageo::alloc(&c->d_fd, 2 * n));
    A(cageo::alloc(&c->d_dist, n));
    A(cageo::alloc(&c->d_lab, n));
    A(cageo::alloc(&c->d_ptr, n));
    A(cageo::alloc(&c->d_bent, 256 * (size_t)c->ntiles));
    if (st == CAGEO_OK &&
        hipHostMalloc(reinterpret_cast<void**>(&c->h_scal), cageo::SC_NSLOTS * sizeof(uint32_t)) !=
            hipSuccess)
        st = CAGEO_ERR_NOMEM;
    // The abort word: host memory the kernels read while they wait at a
    // barrier, so the host can raise it while a launch runs.
    if (st == CAGEO_OK &&
        hipHostMalloc(reinterpret_cast<void**>(&c->h_abort), sizeof(uint32_t),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        st = CAGEO_ERR_NOMEM;
    if (st == CAGEO_OK) {
        __atomic_store_n(c->h_abort, 0u, __ATOMIC_SEQ_CST);
        if (hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_habort), c->h_abort, 0) !=
            hipSuccess)
            st = CAGEO_ERR_HIP;
    }
    if (st == CAGEO_OK) {
        c->bar_timeout_s = cageo::BAR_DEFAULT_TIMEOUT_S;
        st = cageo::barrier_ticks(c->device, c->bar_timeout_s, &c->bar_limit);
    }
    if (st == CAGEO_OK &&
        (hipMemsetAsync(c->d_mark, 0, (size_t)c->ntiles * sizeof(uint32_t), c->stream) != hipSuccess ||
         hipMemsetAsync(c->d_mark_pf, 0, (size_t)c->ntiles * sizeof(uint32_t), c->stream) != hipSuccess ||
         hipMemsetAsync(c->d_ctrl, 0, cageo::CTL_WORDS * sizeof(uint32_t), c->stream) != hipSuccess ||
         hipMemsetAsync(c->d_bar, 0, cageo::BAR_WORDS * sizeof(uint32_t), c->stream) != hipSuccess ||
         hipMemsetAsync(c->d_scal, 0, cageo::SC_NSLOTS * sizeof(uint32_t), c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
        st = CAGEO_ERR_HIP;
    for (hipEvent_t& e : c->ev)
        if (st == CAGEO_OK && hipEventCreate(&e) != hipSuccess) st = CAGEO_ERR_HIP;
    if (st != CAGEO_OK) {
        cageo::free_all(c);
        delete ctx;
        return st;
    }
    *out = ctx;
    return CAGEO_OK;
}

int cageo_context_destroy(cageo_context* ctx) {
    if (!ctx) return CAGEO_ERR_ARG;
    (void)hipStreamSynchronize(ctx->c.stream);
    cageo::free_all(&ctx->c);
    delete ctx;
    return CAGEO_OK;
}

int cageo_set_filter(cageo_context* ctx, int model, int radius, const float* taps) {
    if (!ctx || !taps || radius < 0) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    // The reference wraps a source index once (ConvolutionAlgorithmCPUOpenMP.h:
    // 127-130), which stays in the grid only for radius <= width, height.
    if (radius > c->W || radius > c->H) return CAGEO_ERR_ARG;
    if (model != CAGEO_FILTER_GAUSSIAN && model != CAGEO_FILTER_EXPONENTIAL) return CAGEO_ERR_ARG;
    const int L = 2 * radius + 1;
    std::vector<float> host;
    int fs = L;
    if (model == CAGEO_FILTER_GAUSSIAN) {
        if ((size_t)(cageo::TILE + 2 * radius) * cageo::TILE * sizeof(float) > 160 * 1024)
            return CAGEO_ERR_ARG;
        host.assign(taps, taps + L);
    } else {
        int ty, S;
        size_t lds;
        if (cageo::conv_2d_config(radius, &ty, &S, &lds) != CAGEO_OK) return CAGEO_ERR_ARG;
        fs = (L + 3 + 3) & ~3;
        host.assign((size_t)L * fs, 0.0f);
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) host[(size_t)y * fs + x] = taps[(size_t)y * L + x];
    }
    if (c->d_filter) (void)hipFree(c->d_filter);
    c->d_filter = nullptr;
    CAGEO_TRY(cageo::alloc(&c->d_filter, host.size()));
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_filter, host.data(), host.size() * sizeof(float),
                                 hipMemcpyHostToDevice, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    c->model = model;
    c->radius = radius;
    c->filter_len = (int)host.size();
    if (model == CAGEO_FILTER_EXPONENTIAL) c->h_taps.assign(taps, taps + (size_t)L * L);
    else c->h_taps.clear();
    if (c->conv_fast) CAGEO_TRY(cageo::prepare_low_rank(c));
    return CAGEO_OK;
}

int cageo_context_set_conv_mode(cageo_context* ctx, int mode) {
    if (!ctx || (mode != CAGEO_CONV_EXACT && mode != CAGEO_CONV_FAST)) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    const int old = c->conv_fast;
    c->conv_fast = mode == CAGEO_CONV_FAST;
    if (c->conv_fast && c->model == CAGEO_FILTER_EXPONENTIAL) {
        const int st = cageo::prepare_low_rank(c);
        if (st != CAGEO_OK) {
            c->conv_fast = old;
            return st;
        }
    }
    return CAGEO_OK;
}

int cageo_context_conv_terms(cageo_context* ctx, int* terms) {
    if (!ctx || !terms) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    *terms = (c->conv_fast && c->model == CAGEO_FILTER_EXPONENTIAL) ? (c->fft_ready ? -1 : c->lr_k)
                                                                     : 0;
    return CAGEO_OK;
}

int cageo_random_field(cageo_context* ctx, float* d_field, uint64_t seed, uint32_t realization,
                       uint32_t stream_id, float mean, float std_dev) {
    if (!ctx || !d_field) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::launch_random_field(c, d_field, seed, realization, stream_id, mean, std_dev));
    return cageo::sync(c);
}

int cageo_convolve(cageo_context* ctx, float* d_field) {
    if (!ctx || !d_field) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::launch_convolve(c, d_field, c->d_b));
    return cageo::sync(c);
}

int cageo_realistic_dem(cageo_context* ctx, const float* d_dem, float* d_field, float nodata) {
    if (!ctx || !d_dem || !d_field) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::launch_realistic(c, d_dem, d_field, nodata));
    return cageo::sync(c);
}

int cageo_stream_burn(cageo_context* ctx, float* d_z, const float* d_stream, int bevel_radius,
                      float nodata_dem, float nodata_stream) {
    if (!ctx || !d_z || !d_stream) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::prepare_offsets(c, bevel_radius));
    // scratch: d_st holds the stream flags, d_dist the stream labels
    CAGEO_TRY(cageo::launch_stream_flags(c, d_stream, nodata_stream, c->d_st, c->d_dist));
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_b, d_z, c->n * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->stream));
    CAGEO_TRY(cageo::launch_burn(c, c->d_b, d_z, c->d_st, nodata_dem));
    return cageo::sync(c);
}

int cageo_pit_fill(cageo_context* ctx, float* d_filled, float* d_dem, float nodata) {
    if (!ctx || !d_filled || !d_dem || d_filled == d_dem) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::clear_status(c));
    CAGEO_TRY(cageo::reset_minmax(c));
    CAGEO_TRY(cageo::launch_quantize(c, d_dem, nodata));
    CAGEO_TRY(cageo::pit_fill_quantized(c, d_dem, d_filled, nodata, true));
    return cageo::check_status(c);
}

int cageo_flow_routing(cageo_context* ctx, int8_t* d_flowdir, const float* d_filled,
                       float /*nodata_dem*/) {
    if (!ctx || !d_flowdir || !d_filled) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::clear_status(c));
    CAGEO_TRY(cageo::flow_routing(c, d_filled, d_flowdir));
    return cageo::check_status(c);
}

int cageo_flow_tracing(cageo_context* ctx, int32_t* d_labels, const int8_t* d_flowdir,
                       const float* d_stream, int64_t* long_paths) {
    if (!ctx || !d_labels || !d_flowdir || !d_stream) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    // The trace rule tests stream >= 0 only; the nodata value is irrelevant.
    CAGEO_TRY(cageo::launch_stream_flags(c, d_stream, -__builtin_huge_valf(), c->d_st, c->d_dist));
    CAGEO_TRY(cageo::clear_status(c));
    CAGEO_TRY(cageo::flow_tracing(c, d_flowdir, c->d_st, c->d_dist, d_labels, long_paths));
    return cageo::check_status(c);
}

int cageo_catchment_border(cageo_context* ctx, float* d_mc, const int32_t* d_labels) {
    if (!ctx || !d_mc || !d_labels) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::launch_border(c, d_mc, d_labels));
    return cageo::sync(c);
}

int cageo_to_byte_array(cageo_context* ctx, uint8_t* d_bytes, const float* d_values) {
    if (!ctx || !d_bytes || !d_values) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::launch_to_bytes(c, d_bytes, d_values));
    return cageo::sync(c);
}

int cageo_mc_setup(cageo_context* ctx, const float* d_dem, const float* d_stream,
                   float nodata_dem, float nodata_stream, int bevel_radius, float mean,
                   float std_dev) {
    if (!ctx || !d_dem || !d_stream) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::prepare_offsets(c, bevel_radius));
    CAGEO_TRY(cageo::launch_stream_flags(c, d_stream, nodata_stream, c->d_sflags, c->d_slabel));
    c->dem = d_dem;
    c->streamr = d_stream;
    c->nodata_dem = nodata_dem;
    c->nodata_stream = nodata_stream;
    c->bevel = bevel_radius;
    c->mean = mean;
    c->std_dev = std_dev;
    c->mc_ready = true;
    return cageo::sync(c);
}

int cageo_run_realizations(cageo_context* ctx, float* d_mc, int n, uint64_t seed,
                           uint32_t first) {
    if (!ctx || !d_mc || n < 0) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    if (!c->mc_ready || c->model < 0) return CAGEO_ERR_STATE;
    const bool timed = c->timing;
    if (timed) {
        // Events are recorded on the stream between the stages and read
        // after the last realization, so timing adds no host round trips.
        std::fill(c->stage_ms, c->stage_ms + 7, 0.0f);
        while (c->evpool.size() < (size_t)8 * n) {
            hipEvent_t e;
            CAGEO_HIP_TRY(hipEventCreate(&e));
            c->evpool.push_back(e);
        }
    }
    CAGEO_TRY(cageo::clear_status(c));
    for (int r = 0; r < n; ++r) {
        hipEvent_t* evs = timed ? &c->evpool[(size_t)8 * r] : nullptr;
        if (timed) CAGEO_HIP_TRY(hipEventRecord(evs[0], c->stream));
        // RandomFieldAlgorithm::execute(rand, seed), main.cpp:189-192
        CAGEO_TRY(cageo::launch_random_field(c, c->d_a, seed, first + (uint32_t)r, 0u, c->mean,
                                             c->std_dev));
        if (timed) CAGEO_HIP_TRY(hipEventRecord(evs[1], c->stream));
        CAGEO_TRY(cageo::realization_body(c, d_mc, evs));
    }
    const int st = cageo::check_status(c);
    if (timed && st == CAGEO_OK) {
        for (int r = 0; r < n; ++r)
            for (int s = 0; s < 7; ++s) {
                float ms = 0.0f;
                CAGEO_HIP_TRY(hipEventElapsedTime(&ms, c->evpool[(size_t)8 * r + s],
                                                  c->evpool[(size_t)8 * r + s + 1]));
                c->stage_ms[s] += ms;
            }
    }
    return st;
}

int cageo_realization_from_noise(cageo_context* ctx, float* d_noise, float* d_mc,
                                 float* d_filled_out, int8_t* d_flowdir_out,
                                 int32_t* d_labels_out) {
    if (!ctx || !d_noise) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    if (!c->mc_ready || c->model < 0) return CAGEO_ERR_STATE;
    CAGEO_HIP_TRY(hipMemcpyAsync(c->d_a, d_noise, c->n * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->stream));
    CAGEO_TRY(cageo::clear_status(c));
    CAGEO_TRY(cageo::realization_body(c, d_mc, nullptr));
    CAGEO_TRY(cageo::check_status(c));
    if (d_filled_out)
        CAGEO_HIP_TRY(hipMemcpyAsync(d_filled_out, c->d_w, c->n * sizeof(float),
                                     hipMemcpyDeviceToDevice, c->stream));
    if (d_flowdir_out)
        CAGEO_HIP_TRY(hipMemcpyAsync(d_flowdir_out, c->d_fd, 2 * c->n, hipMemcpyDeviceToDevice,
                                     c->stream));
    if (d_labels_out)
        CAGEO_HIP_TRY(hipMemcpyAsync(d_labels_out, c->d_lab, c->n * sizeof(int32_t),
                                     hipMemcpyDeviceToDevice, c->stream));
    return cageo::sync(c);
}

int cageo_iteration_counts(cageo_context* ctx, int32_t* out) {
    if (!ctx || !out) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    uint32_t h[cageo::CTL_WORDS];
    CAGEO_HIP_TRY(hipMemcpyAsync(h, c->d_ctrl, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(c->stream));
    out[0] = (int32_t)h[cageo::CTL_ROUNDS];
    out[1] = (int32_t)h[cageo::CTL_PASSES];
    out[2] = (int32_t)h[cageo::CTL_FR_PASSES];
    out[3] = (int32_t)h[cageo::CTL_TR_ROUNDS];
    return CAGEO_OK;
}

int cageo_context_configure(cageo_context* ctx, int persistent_workgroups, int conv_rows) {
    if (!ctx || persistent_workgroups < 0) return CAGEO_ERR_ARG;
    if (conv_rows != 0 && conv_rows != 4 && conv_rows != 8 && conv_rows != 16) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    int caps[3];
    CAGEO_TRY(cageo::pit_fill_grid_cap(c, &caps[0]));
    CAGEO_TRY(cageo::flow_routing_grid_cap(c, &caps[1]));
    CAGEO_TRY(cageo::flow_tracing_grid_cap(c, &caps[2]));
    const int cap = std::min(caps[0], std::min(caps[1], caps[2]));
    if (persistent_workgroups > cap) return CAGEO_ERR_ARG;
    if (persistent_workgroups == 0) {
        c->pf_grid = caps[0];
        c->fr_grid = caps[1];
        c->tr_grid = caps[2];
    } else {
        c->pf_grid = c->fr_grid = c->tr_grid = persistent_workgroups;
    }
    c->conv_ty = conv_rows;
    return CAGEO_OK;
}

int cageo_context_long_paths(cageo_context* ctx, int64_t* out) {
    if (!ctx || !out) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    CAGEO_TRY(cageo::read_scalars(c));
    *out = (int64_t)c->h_scal[cageo::SC_LONGRUN];
    return CAGEO_OK;
}

int cageo_context_abort(cageo_context* ctx, int set) {
    if (!ctx || !ctx->c.h_abort) return CAGEO_ERR_ARG;
    __atomic_store_n(ctx->c.h_abort, set ? 1u : 0u, __ATOMIC_SEQ_CST);
    return CAGEO_OK;
}

int cageo_context_set_barrier_timeout(cageo_context* ctx, double seconds) {
    if (!ctx || !(seconds >= 0.0)) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    uint64_t t = 0;
    CAGEO_TRY(cageo::barrier_ticks(c->device, seconds, &t));
    c->bar_limit = t;
    c->bar_timeout_s = seconds;
    return CAGEO_OK;
}

int cageo_stage_times(cageo_context* ctx, float* out, int enable) {
    if (!ctx) return CAGEO_ERR_ARG;
    Context* c = &ctx->c;
    if (out) std::memcpy(out, c->stage_ms, sizeof(c->stage_ms));
    c->timing = enable != 0;
    return CAGEO_OK;
}

}  // extern "C"
