// Random field stage: RandomFieldAlgorithm<T>::execute, RandomFieldAlgorithm.h:
// 82-93 (reference GPU kernels RandomFieldAlgorithmGPUCUDA.cu:30-88, CPU
// RandomFieldAlgorithmCPUOpenMP.h:47-76).  value = mean + r * std in float as
// in RandomFieldAlgorithmCPUOpenMP.h:72, r from the counter-based generator
// in rng.hpp: no per-thread state, no setup kernel, every cell independent.
#include "common.hpp"
#include "rng.hpp"

namespace cageo {

// `realization_base`, when not null, is added to `realization` (a graph
// replayed once per realization reads its index from device memory).
__global__ __launch_bounds__(256) void k_random_field(float* __restrict__ out, int64_t n,
                                                      uint64_t seed, uint32_t realization,
                                                      const uint32_t* __restrict__ realization_base,
                                                      uint32_t stream_id, float mean,
                                                      float std_dev, int vec4) {
    if (realization_base) realization += *realization_base;
    const int64_t nblocks = (n + 3) >> 2;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks;
         b += (int64_t)gridDim.x * blockDim.x) {
        float g[4];
        normal4((uint64_t)b, realization, stream_id, seed, g);
        float v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float scaled = g[k] * std_dev;
            v[k] = mean + scaled;
        }
        const int64_t i = b << 2;
        if (vec4 && i + 3 < n) {
            *reinterpret_cast<float4*>(out + i) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (i + k < n) out[i + k] = v[k];
        }
    }
}

int launch_random_field(Context* c, float* out, uint64_t seed, uint32_t realization,
                        uint32_t stream_id, float mean, float std_dev,
                        const uint32_t* realization_base) {
    const int64_t nblocks = (int64_t)((c->n + 3) / 4);
    int grid = (int)((nblocks + 255) / 256);
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    const int vec4 = ((uintptr_t)out % 16) == 0;
    hipLaunchKernelGGL(k_random_field, dim3(grid), dim3(256), 0, c->stream, out,
                       (int64_t)c->n, seed, realization, realization_base, stream_id, mean, std_dev,
                       vec4);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo
