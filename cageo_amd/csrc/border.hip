// Catchment border + Monte Carlo accumulation in one pass:
// CatchmentBorderAlgorithm<T,U>::execute, CatchmentBorderAlgorithm.h:53-74,
// with performExtractDrainageBorder, CatchmentBorderAlgorithmCPUOpenMP.h:
// 30-82, and the AddSurfaces(noData = -1) that adds the border raster into
// the counter (:60-69).  A cell is a border cell when its label is not LAKE
// and some in-grid 8-neighbour has a larger label that is not LAKE; the
// counter takes mc + border unless mc == -1.
#include "common.hpp"

namespace cageo {

namespace {

__device__ constexpr int BDX[8] = {0, 1, 1, 1, 0, -1, -1, -1};
__device__ constexpr int BDY[8] = {-1, -1, 0, 1, 1, 1, 0, -1};

__global__ __launch_bounds__(256) void k_border(float* __restrict__ mc,
                                                const int32_t* __restrict__ lab, int W,
                                                int H) {
    // Blocks go round robin over the 8 XCDs; renumbering them so that each
    // XCD takes a contiguous band of tiles keeps a tile's neighbours, whose
    // labels it reads at its edges, in the same L2 (a bijection on the
    // block numbers when their count is a multiple of 8).
    const unsigned nb = gridDim.x * gridDim.y;
    unsigned t = blockIdx.y * gridDim.x + blockIdx.x;
    if ((nb & 7) == 0) t = (t & 7) * (nb >> 3) + (t >> 3);
    const int x = (int)(t % gridDim.x) * 64 + (threadIdx.x & 63);
    const int y = (int)(t / gridDim.x) * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const size_t i = (size_t)y * W + x;
    const int32_t c = lab[i];
    const float m = mc[i];
    // all neighbour labels are loaded before any is tested (no dependent
    // loads); an off-grid neighbour reads as LAKE, which never qualifies
    int32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int nx = x + BDX[k], ny = y + BDY[k];
        v[k] = (nx < 0 || ny < 0 || nx >= W || ny >= H) ? CAGEO_FLOWTRACE_LAKE
                                                         : lab[(size_t)ny * W + nx];
    }
    bool any = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) any |= c < v[k] && v[k] != CAGEO_FLOWTRACE_LAKE;
    const float b = (c != CAGEO_FLOWTRACE_LAKE && any) ? 1.0f : 0.0f;
    if (m != -1.0f) mc[i] = m + b;
}

}  // namespace

int launch_border(Context* c, float* mc, const int32_t* lab) {
    dim3 grid((c->W + 63) / 64, (c->H + 3) / 4);
    hipLaunchKernelGGL(k_border, grid, dim3(256), 0, c->stream, mc, lab, c->W, c->H);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
