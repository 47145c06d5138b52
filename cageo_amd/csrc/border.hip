// Catchment border + Monte Carlo accumulation in one pass:
// CatchmentBorderAlgorithm<T,U>::execute, CatchmentBorderAlgorithm.h:53-74,
// with performExtractDrainageBorder, CatchmentBorderAlgorithmCPUOpenMP.h:
// 30-82, and the AddSurfaces(noData = -1) that adds the border raster into
// the counter (:60-69).  A cell is a border cell when its label is not LAKE
// and some in-grid 8-neighbour has a larger label that is not LAKE; the
// counter takes mc + border unless mc == -1.
#include "tiles.hpp"

namespace cageo {

namespace {

// One workgroup per 64 x 64 block, each wave a strip of BR_ROWS rows: lane
// x walks its column with one label load per row (rows y0-1 .. y0+BR_ROWS),
// the labels of columns x-1 and x+1 come from the neighbouring lanes, and
// only lanes 0 and 63 load the block's outer columns.  Every load of the
// strip is issued before any is used.  Off-grid labels read as LAKE, which
// never qualifies.
constexpr int BR_ROWS = 16;

__global__ __launch_bounds__(256) void k_border(float* __restrict__ mc,
                                                const int32_t* __restrict__ lab, int W,
                                                int H) {
    // Blocks go round robin over the 8 XCDs; renumbering them so that each
    // XCD takes a contiguous band of blocks keeps a block's neighbours, whose
    // labels it reads at its edges, in the same L2 (a bijection on the
    // block numbers when their count is a multiple of 8).
    const unsigned nb = gridDim.x * gridDim.y;
    unsigned t = blockIdx.y * gridDim.x + blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int lane = threadIdx.x & 63;
    const int x = (int)(t % gridDim.x) * 64 + lane;
    const int y0 = (int)(t / gridDim.x) * 64 + (int)(threadIdx.x >> 6) * BR_ROWS;
    const bool edge = lane == 0 || lane == 63;
    const int ex = lane == 0 ? x - 1 : x + 1;  // the outer column of an edge lane
    int32_t cv[BR_ROWS + 2], ev[BR_ROWS + 2];
    float m[BR_ROWS];
#pragma unroll
    for (int r = 0; r < BR_ROWS + 2; ++r) {
        const int y = y0 - 1 + r;
        const bool yin = y >= 0 && y < H;
        const bool cin = yin && x < W, ein = yin && edge && ex >= 0 && ex < W;
        cv[r] = lab[cin ? (size_t)y * W + x : 0];
        ev[r] = lab[ein ? (size_t)y * W + ex : 0];
        if (!cin) cv[r] = CAGEO_FLOWTRACE_LAKE;
        if (!ein) ev[r] = CAGEO_FLOWTRACE_LAKE;
    }
#pragma unroll
    for (int k = 0; k < BR_ROWS; ++k) {
        const int y = y0 + k;
        m[k] = mc[(x < W && y < H) ? (size_t)y * W + x : 0];
    }
    // label of column x-1 / x+1 in strip row r
    auto left = [&](int r) { const int v = lane_prev(cv[r]); return lane == 0 ? ev[r] : v; };
    auto right = [&](int r) { const int v = lane_next(cv[r]); return lane == 63 ? ev[r] : v; };
    auto nb_ok = [](int32_t c, int32_t v) { return c < v && v != CAGEO_FLOWTRACE_LAKE; };
    int32_t l[BR_ROWS + 2], rr[BR_ROWS + 2];
#pragma unroll
    for (int r = 0; r < BR_ROWS + 2; ++r) {
        l[r] = left(r);
        rr[r] = right(r);
    }
#pragma unroll
    for (int k = 0; k < BR_ROWS; ++k) {
        const int y = y0 + k, r = k + 1;
        const int32_t c = cv[r];
        const bool any = nb_ok(c, cv[r - 1]) || nb_ok(c, rr[r - 1]) || nb_ok(c, rr[r]) ||
                         nb_ok(c, rr[r + 1]) || nb_ok(c, cv[r + 1]) || nb_ok(c, l[r + 1]) ||
                         nb_ok(c, l[r]) || nb_ok(c, l[r - 1]);
        const float b = (c != CAGEO_FLOWTRACE_LAKE && any) ? 1.0f : 0.0f;
        if (x < W && y < H && m[k] != -1.0f) mc[(size_t)y * W + x] = m[k] + b;
    }
}

}  // namespace

int launch_border(Context* c, float* mc, const int32_t* lab) {
    dim3 grid((c->W + 63) / 64, (c->H + 63) / 64);
    hipLaunchKernelGGL(k_border, grid, dim3(256), 0, c->stream, mc, lab, c->W, c->H);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
