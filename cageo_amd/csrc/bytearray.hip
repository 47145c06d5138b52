// Visualisation: a float raster scaled to bytes for the JPEG writer, the
// device counterpart of ToByteArrayAlgorithm (ToByteArrayAlgorithmGPUCUDA.cu:
// 22-84).  min and max are taken over every cell (nodata counters included,
// as thrust::minmax_element does there) and the scale divides by max, not by
// max - min, exactly as the reference does.
#include "common.hpp"

namespace cageo {

namespace {

__global__ __launch_bounds__(256) void k_minmax_raw(const float* __restrict__ in, int64_t n,
                                                    uint32_t* __restrict__ scal) {
    uint32_t mn = 0xffffffffu, mx = 0u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = float_to_ordered(in[i]);
        mn = min(mn, o);
        mx = max(mx, o);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    }
    __shared__ uint32_t s_mn, s_mx;
    if (threadIdx.x == 0) {
        s_mn = 0xffffffffu;
        s_mx = 0u;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&s_mn, mn);
        atomicMax(&s_mx, mx);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(&scal[SC_MIN], s_mn);
        atomicMax(&scal[SC_MAX], s_mx);
    }
}

// Byte_t = (Byte_t) rintf(255 * ((v - min) / max)).  Values outside 0..255
// (a negative min) wrap modulo 256 through an int conversion, and 0/0 gives
// 0, which is what the x86 float->int->byte conversion of the CPU path does.
__global__ __launch_bounds__(256) void k_to_bytes(uint8_t* __restrict__ out,
                                                  const float* __restrict__ in, int64_t n,
                                                  const uint32_t* __restrict__ scal) {
    const float mn = ordered_to_float(scal[SC_MIN]);
    const float mx = ordered_to_float(scal[SC_MAX]);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = rintf(255.0f * ((in[i] - mn) / mx));
        int32_t q = 0;
        if (v == v && v > -2147483648.0f && v < 2147483648.0f) q = (int32_t)v;
        out[i] = (uint8_t)q;
    }
}

}  // namespace

int launch_to_bytes(Context* c, uint8_t* out, const float* in) {
    CAGEO_HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(c->d_scal + SC_MIN),
                                    0xffffffffu, 1, c->stream));
    CAGEO_HIP_TRY(hipMemsetAsync(c->d_scal + SC_MAX, 0, sizeof(uint32_t), c->stream));
    const int64_t n = (int64_t)c->n;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_minmax_raw, dim3(grid), dim3(256), 0, c->stream, in, n, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_to_bytes, dim3(grid), dim3(256), 0, c->stream, out, in, n, c->d_scal);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
