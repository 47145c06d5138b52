// Border-probability surrogate: inference and its Monte Carlo check behind
// the C ABI, on the library's BF16 MFMA kernels (nn.hip).
//
// No reference counterpart: the reference only runs the Monte Carlo.  The
// network (a 4-level U-Net, base b channels) maps the z-scored DEM, the
// stream mask, the error RMSE / 3 m and the practical range / 200 m to the
// per-cell probability that a realization puts a catchment border there.
// Layers (3x3 convolutions, ReLU, 2x2 average pooling, nearest 2x
// upsampling folded into the decoder's convolution, skip concatenation):
//   enc1 32->b->b, enc2 b->2b->2b, enc3 2b->4b->4b, mid 4b->8b->8b,
//   dec3 [8b up | 4b]->4b->4b, dec2 [4b up | 2b]->2b->2b,
//   dec1 [2b up | b]->b->b, head 1x1 b->1 with a sigmoid.
// The same network is trained by cageo_amd/surrogate.py on these kernels;
// weights come from its `.bin` files.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

extern "C" {
int cageo_nn_conv3x3(void* stream, const uint16_t* in1, int c1, int up1, const uint16_t* in2,
                     int c2, const uint16_t* w, const float* bias, uint16_t* out, float* outf,
                     int n, int h, int wdt, int cout, int relu);
int cageo_nn_avgpool2(void* stream, const uint16_t* in, uint16_t* out, int n, int h, int w, int c);
int cageo_nn_head(void* stream, const uint16_t* act, const float* w, const float* b, float* logit,
                  long long p, int c, int sigmoid);
int cageo_nn_conv3x3_fusable(int h, int w, int cout, int head);
int cageo_nn_conv3x3_fused(void* stream, const uint16_t* in1, int c1, int up1, const uint16_t* in2,
                           int c2, const uint16_t* w, const float* bias, uint16_t* out,
                           uint16_t* pool, const float* head_w, const float* head_b, float* prob,
                           int head_sigmoid, int n, int h, int wdt, int cout, int relu);
}

namespace cageo {

namespace {

constexpr int IN_PAD = 32;  // input channels the first convolution reads (4 used)

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Partial sums (count, sum, sum of squares) of the valid DEM cells, one
// triple per workgroup, in double.
__global__ void k_dem_partials(const float* __restrict__ dem, long long n, float nodata,
                               double* __restrict__ part) {
    __shared__ double s[3][256];
    double c = 0.0, a = 0.0, q = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float v = dem[i];
        if (v != nodata) {
            c += 1.0;
            a += v;
            q += (double)v * v;
        }
    }
    s[0][threadIdx.x] = c;
    s[1][threadIdx.x] = a;
    s[2][threadIdx.x] = q;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[(size_t)blockIdx.x * 3 + k] = s[k][0];
}

// Mean and (population) standard deviation from the partials -> stats[0..1].
// One wave of 64 threads: strided partial sums, then a butterfly.
__global__ void k_dem_stats(const double* __restrict__ part, int nparts, float* __restrict__ stats) {
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    double c = 0.0, a = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) {
        c += part[(size_t)i * 3];
        a += part[(size_t)i * 3 + 1];
        q += part[(size_t)i * 3 + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off);
        a += __shfl_xor(a, off);
        q += __shfl_xor(q, off);
    }
    if (threadIdx.x != 0) return;
    double mu = c > 0 ? a / c : 0.0;
    double var = c > 0 ? q / c - mu * mu : 1.0;
    double sd = var > 0 ? sqrt(var) : 0.0;
    stats[0] = (float)mu;
    stats[1] = (float)(sd > 1e-6 ? sd : 1e-6);
}

// Input planes, NHWC bf16 with IN_PAD channels over the padded hp x wp grid
// (zero outside the W x H raster): z-scored DEM (nodata 0), stream mask,
// rmse / 3, practical range / 200.
__global__ void k_make_inputs(const float* __restrict__ dem, const float* __restrict__ stream,
                              int W, int H, int wp, int hp, float nodata_dem, float nodata_stream,
                              float c2, float c3, const float* __restrict__ stats,
                              uint16_t* __restrict__ x) {
    const float mu = stats[0], sd = stats[1];
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < (long long)wp * hp;
         p += (long long)gridDim.x * blockDim.x) {
        const int yy = (int)(p / wp), xx = (int)(p - (long long)yy * wp);
        uint4 v0 = make_uint4(0, 0, 0, 0);
        if (yy < H && xx < W) {
            const size_t i = (size_t)yy * W + xx;
            const float d = dem[i];
            const float z = d != nodata_dem ? (d - mu) / sd : 0.0f;
            const float s = stream[i] != nodata_stream ? 1.0f : 0.0f;
            v0.x = (uint32_t)to_bf16(z) | ((uint32_t)to_bf16(s) << 16);
            v0.y = (uint32_t)to_bf16(c2) | ((uint32_t)to_bf16(c3) << 16);
        }
        uint4* o = reinterpret_cast<uint4*>(x + (size_t)p * IN_PAD);
        o[0] = v0;
        o[1] = make_uint4(0, 0, 0, 0);
        o[2] = make_uint4(0, 0, 0, 0);
        o[3] = make_uint4(0, 0, 0, 0);
    }
}

// The first convolution (enc1a, base 32) straight from the DEM and the
// stream raster, for cageo_surrogate_predict: the input planes of
// k_make_inputs are formed in LDS for a 4-row x 64-pixel tile and its halo
// and never stored, and the convolution reads only the 4 planes it uses:
// K = 9 taps x 4 channels in two 32-wide MFMA steps (taps 0-7, then tap 8
// and zeros) instead of 9 taps x 32 padded channels.  Same products, fp32
// accumulation in another order than the padded layer's.  Output
// [hp][wp][32] bf16 after bias and ReLU, staged in LDS and written with
// 16-byte stores.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int FL_TX = 64, FL_TY = 4, FL_AX = FL_TX + 2, FL_AN = (FL_TY + 2) * FL_AX;
constexpr int FL_PS = 40;  // staged pixel stride (bf16): 32 channels + 16 bytes

__global__ __launch_bounds__(256) void k_first_layer(
    const float* __restrict__ dem, const float* __restrict__ stream, int W, int H, int wp, int hp,
    float nodata_dem, float nodata_stream, float c2, float c3, const float* __restrict__ stats,
    const uint16_t* __restrict__ w, const float* __restrict__ bias, uint16_t* __restrict__ out) {
    __shared__ uint2 pl[FL_AN];  // the 4 bf16 planes of each halo-tile pixel
    __shared__ __align__(16) uint16_t so[FL_TY * FL_TX * FL_PS];
    const int x0 = blockIdx.x * FL_TX, y0 = blockIdx.y * FL_TY;
    const float mu = stats[0], sd = stats[1];
    const uint32_t cc = (uint32_t)to_bf16(c2) | ((uint32_t)to_bf16(c3) << 16);
    {
        // all loads first (a pixel outside the raster loads cell 0), then
        // the planes; zero outside the raster, as k_make_inputs writes them
        constexpr int IT = (FL_AN + 255) / 256;
        float dv[IT], sv[IT];
        bool in[IT];
#pragma unroll
        for (int j = 0; j < IT; ++j) {
            const int i = threadIdx.x + 256 * j;
            const int yy = y0 - 1 + i / FL_AX, xx = x0 - 1 + i % FL_AX;
            in[j] = i < FL_AN && yy >= 0 && xx >= 0 && yy < H && xx < W;
            const size_t k = in[j] ? (size_t)yy * W + xx : 0;
            dv[j] = dem[k];
            sv[j] = stream[k];
        }
#pragma unroll
        for (int j = 0; j < IT; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (i >= FL_AN) continue;
            uint2 v = make_uint2(0u, 0u);
            if (in[j]) {
                const float z = dv[j] != nodata_dem ? (dv[j] - mu) / sd : 0.0f;
                const float sm = sv[j] != nodata_stream ? 1.0f : 0.0f;
                v = make_uint2((uint32_t)to_bf16(z) | ((uint32_t)to_bf16(sm) << 16), cc);
            }
            pl[i] = v;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kb = lane >> 4;
    // B: lane (column row, K block kb) holds K 8kb..8kb+7 = taps 2kb, 2kb+1 x
    // channels 0..3 (first step), tap 8 x channels 0..3 for kb 0 (second)
    bf16x8 b0[2], b1[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int co = 16 * q + row;
        const uint2 t0 = *reinterpret_cast<const uint2*>(w + ((2 * kb) * 32 + co) * 32);
        const uint2 t1 = *reinterpret_cast<const uint2*>(w + ((2 * kb + 1) * 32 + co) * 32);
        const uint2 t8 = kb == 0 ? *reinterpret_cast<const uint2*>(w + (8 * 32 + co) * 32)
                                 : make_uint2(0u, 0u);
        b0[q] = __builtin_bit_cast(bf16x8, make_uint4(t0.x, t0.y, t1.x, t1.y));
        b1[q] = __builtin_bit_cast(bf16x8, make_uint4(t8.x, t8.y, 0u, 0u));
    }
    __syncthreads();
    f32x4 acc[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int px = 16 * m + row;  // pixel of the wave's row; halo-shifted taps below
        const int ta = 2 * kb, tb = 2 * kb + 1;
        const uint2 a0 = pl[(wave + ta / 3) * FL_AX + px + ta % 3];
        const uint2 a1 = pl[(wave + tb / 3) * FL_AX + px + tb % 3];
        const uint2 a8 = kb == 0 ? pl[(wave + 2) * FL_AX + px + 2] : make_uint2(0u, 0u);
        const bf16x8 A0 = __builtin_bit_cast(bf16x8, make_uint4(a0.x, a0.y, a1.x, a1.y));
        const bf16x8 A1 = __builtin_bit_cast(bf16x8, make_uint4(a8.x, a8.y, 0u, 0u));
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, b0[q], f32x4{0.f, 0.f, 0.f, 0.f},
                                                                0, 0, 0);
            acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, b1[q], acc[m][q], 0, 0, 0);
        }
    }
    // accumulator row 4 kb + r = pixel, column = channel 16 q + row
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = 16 * q + row;
        const float bc = bias[c];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                so[(wave * FL_TX + 16 * m + 4 * kb + r) * FL_PS + c] =
                    to_bf16(fmaxf(acc[m][q][r] + bc, 0.0f));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < FL_TY * FL_TX * 4; i += 256) {
        const int px = i >> 2, part = i & 3;
        const int y = y0 + px / FL_TX, x = x0 + px % FL_TX;
        if (x < wp && y < hp)
            *reinterpret_cast<uint4*>(out + ((size_t)y * wp + x) * 32 + 8 * part) =
                *reinterpret_cast<const uint4*>(so + px * FL_PS + 8 * part);
    }
}

// Crop of the padded probability map.
__global__ void k_crop(const float* __restrict__ in, int wp, float* __restrict__ out, int W, int H) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)W * H;
         i += (long long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long long)y * W);
        out[i] = in[(size_t)y * wp + x];
    }
}

constexpr int RED_BLOCKS = 1024;

struct Layer {
    const char* name;
    int level;  // resolution 1/2^level
    int up;     // first input is the coarser decoder tensor (upsampled)
};
// enc / mid / dec layers in execution order (names of the weights file)
constexpr Layer LAYERS[] = {
    {"enc1a", 0, 0}, {"enc1b", 0, 0}, {"enc2a", 1, 0}, {"enc2b", 1, 0}, {"enc3a", 2, 0},
    {"enc3b", 2, 0}, {"mida", 3, 0},  {"midb", 3, 0},  {"dec3a", 2, 1}, {"dec3b", 2, 0},
    {"dec2a", 1, 1}, {"dec2b", 1, 0}, {"dec1a", 0, 1}, {"dec1b", 0, 0}};

}  // namespace

}  // namespace cageo

struct cageo_surrogate {
    hipStream_t stream = nullptr;
    int base = 0;
    std::map<std::string, void*> dev;           // weights on the device
    std::map<std::string, std::vector<int>> shape;
    int hp = 0, wp = 0;                         // workspace for this padded size
    std::vector<uint16_t*> act;                 // activations, see forward()
    float* prob_pad = nullptr;
    double* part = nullptr;
    float* stats = nullptr;
};

namespace cageo {

namespace {

void free_workspace(cageo_surrogate* s) {
    for (uint16_t* p : s->act)
        if (p) (void)hipFree(p);
    s->act.clear();
    if (s->prob_pad) (void)hipFree(s->prob_pad);
    s->prob_pad = nullptr;
    s->hp = s->wp = 0;
}

// Activation buffers for an hp x wp input (multiples of 8): index 0 is the
// input x; per level 4 buffers (a temporary, the skip tensor, the decoder
// output, the pooled input of the next level) sized for the level's widest
// tensor.
int ensure_workspace(cageo_surrogate* s, int hp, int wp) {
    if (s->hp == hp && s->wp == wp) return CAGEO_OK;
    free_workspace(s);
    const int b = s->base;
    const size_t px0 = (size_t)hp * wp;
    auto A = [&](size_t elems) -> int {
        uint16_t* p = nullptr;
        if (hipMalloc(reinterpret_cast<void**>(&p), elems * sizeof(uint16_t)) != hipSuccess)
            return CAGEO_ERR_NOMEM;
        s->act.push_back(p);
        return CAGEO_OK;
    };
    CAGEO_TRY(A(px0 * IN_PAD));  // x
    for (int lv = 0; lv < 4; ++lv) {
        const size_t px = px0 >> (2 * lv);
        const int c = b << lv;
        CAGEO_TRY(A(px * c));      // temporary (first conv of the level)
        CAGEO_TRY(A(px * c));      // skip / encoder output
        CAGEO_TRY(A(px * c));      // decoder output
        CAGEO_TRY(A(px / 4 * c));  // pooled input of the next level (c channels at 1/4 the pixels)
    }
    if (hipMalloc(reinterpret_cast<void**>(&s->prob_pad), px0 * sizeof(float)) != hipSuccess)
        return CAGEO_ERR_NOMEM;
    s->hp = hp;
    s->wp = wp;
    return CAGEO_OK;
}

int conv(cageo_surrogate* s, const char* name, const uint16_t* in1, int c1, int up1,
         const uint16_t* in2, int c2, uint16_t* out, int h, int w) {
    const std::string n(name);
    auto wi = s->dev.find(n + ".w"), bi = s->dev.find(n + ".b");
    if (wi == s->dev.end() || bi == s->dev.end()) return CAGEO_ERR_STATE;
    const int cout = s->shape[n + ".b"][0];
    return cageo_nn_conv3x3(s->stream, in1, c1, up1, in2, c2,
                            static_cast<const uint16_t*>(wi->second),
                            static_cast<const float*>(bi->second), out, nullptr, 1, h, w, cout, 1);
}

// A level's second convolution, its output also 2x2-average pooled: in one
// launch where the staged convolution applies (the same bits as the pool
// kernel after it), else the two kernels.
int conv_pool(cageo_surrogate* s, const char* name, const uint16_t* in1, int c1, uint16_t* out,
              uint16_t* pooled, int h, int w) {
    const std::string n(name);
    auto wi = s->dev.find(n + ".w"), bi = s->dev.find(n + ".b");
    if (wi == s->dev.end() || bi == s->dev.end()) return CAGEO_ERR_STATE;
    const int cout = s->shape[n + ".b"][0];
    if (cageo_nn_conv3x3_fusable(h, w, cout, 0))
        return cageo_nn_conv3x3_fused(s->stream, in1, c1, 0, nullptr, 0,
                                      static_cast<const uint16_t*>(wi->second),
                                      static_cast<const float*>(bi->second), out, pooled, nullptr,
                                      nullptr, nullptr, 0, 1, h, w, cout, 1);
    CAGEO_TRY(conv(s, name, in1, c1, 0, nullptr, 0, out, h, w));
    return cageo_nn_avgpool2(s->stream, out, pooled, 1, h, w, cout);
}

// The last convolution and the sigmoid head into prob_pad: one launch where
// the staged convolution holds all of its channels in a workgroup (its
// output then never reaches memory), else the convolution into `tmp` and
// the head kernel.
int conv_head(cageo_surrogate* s, const char* name, const uint16_t* in1, int c1, uint16_t* tmp,
              int h, int w) {
    const std::string n(name);
    auto wi = s->dev.find(n + ".w"), bi = s->dev.find(n + ".b");
    if (wi == s->dev.end() || bi == s->dev.end()) return CAGEO_ERR_STATE;
    const int cout = s->shape[n + ".b"][0];
    const float* hw = static_cast<const float*>(s->dev["head.w"]);
    const float* hb = static_cast<const float*>(s->dev["head.b"]);
    if (cageo_nn_conv3x3_fusable(h, w, cout, 1))
        return cageo_nn_conv3x3_fused(s->stream, in1, c1, 0, nullptr, 0,
                                      static_cast<const uint16_t*>(wi->second),
                                      static_cast<const float*>(bi->second), nullptr, nullptr, hw,
                                      hb, s->prob_pad, 1, 1, h, w, cout, 1);
    CAGEO_TRY(conv(s, name, in1, c1, 0, nullptr, 0, tmp, h, w));
    return cageo_nn_head(s->stream, tmp, hw, hb, s->prob_pad, (long long)h * w, cout, 1);
}

// Forward pass of x (bf16 [hp][wp][IN_PAD] = act[0]) into prob_pad; with
// `first_done`, enc1a's output is already in its buffer (k_first_layer).
int forward(cageo_surrogate* s, bool first_done = false) {
    const int b = s->base, hp = s->hp, wp = s->wp;
    auto T = [&](int lv) { return s->act[1 + 4 * lv + 0]; };  // temporary
    auto E = [&](int lv) { return s->act[1 + 4 * lv + 1]; };  // encoder output (skip)
    auto D = [&](int lv) { return s->act[1 + 4 * lv + 2]; };  // decoder output
    auto P = [&](int lv) { return s->act[1 + 4 * lv + 3]; };  // pooled input of level lv + 1
    auto H = [&](int lv) { return hp >> lv; };
    auto Wd = [&](int lv) { return wp >> lv; };
    const uint16_t* x = s->act[0];
    // encoder
    if (!first_done) CAGEO_TRY(conv(s, "enc1a", x, IN_PAD, 0, nullptr, 0, T(0), H(0), Wd(0)));
    CAGEO_TRY(conv_pool(s, "enc1b", T(0), b, E(0), P(0), H(0), Wd(0)));
    CAGEO_TRY(conv(s, "enc2a", P(0), b, 0, nullptr, 0, T(1), H(1), Wd(1)));
    CAGEO_TRY(conv_pool(s, "enc2b", T(1), 2 * b, E(1), P(1), H(1), Wd(1)));
    CAGEO_TRY(conv(s, "enc3a", P(1), 2 * b, 0, nullptr, 0, T(2), H(2), Wd(2)));
    CAGEO_TRY(conv_pool(s, "enc3b", T(2), 4 * b, E(2), P(2), H(2), Wd(2)));
    CAGEO_TRY(conv(s, "mida", P(2), 4 * b, 0, nullptr, 0, T(3), H(3), Wd(3)));
    CAGEO_TRY(conv(s, "midb", T(3), 8 * b, 0, nullptr, 0, E(3), H(3), Wd(3)));
    // decoder: the coarser tensor is upsampled inside the convolution's loads
    CAGEO_TRY(conv(s, "dec3a", E(3), 8 * b, 1, E(2), 4 * b, T(2), H(2), Wd(2)));
    CAGEO_TRY(conv(s, "dec3b", T(2), 4 * b, 0, nullptr, 0, D(2), H(2), Wd(2)));
    CAGEO_TRY(conv(s, "dec2a", D(2), 4 * b, 1, E(1), 2 * b, T(1), H(1), Wd(1)));
    CAGEO_TRY(conv(s, "dec2b", T(1), 2 * b, 0, nullptr, 0, D(1), H(1), Wd(1)));
    CAGEO_TRY(conv(s, "dec1a", D(1), 2 * b, 1, E(0), b, T(0), H(0), Wd(0)));
    return conv_head(s, "dec1b", T(0), b, D(0), H(0), Wd(0));
}

bool read_exact(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

}  // namespace

}  // namespace cageo

using cageo::Context;

extern "C" {

int cageo_surrogate_destroy(cageo_surrogate* s) {
    if (!s) return CAGEO_ERR_ARG;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    cageo::free_workspace(s);
    for (auto& kv : s->dev) (void)hipFree(kv.second);
    if (s->part) (void)hipFree(s->part);
    if (s->stats) (void)hipFree(s->stats);
    delete s;
    return CAGEO_OK;
}

int cageo_surrogate_create(cageo_surrogate** out, const char* weights_path, void* stream) {
    if (!out || !weights_path) return CAGEO_ERR_ARG;
    FILE* f = std::fopen(weights_path, "rb");
    if (!f) return CAGEO_ERR_ARG;
    cageo_surrogate* s = new cageo_surrogate();
    s->stream = reinterpret_cast<hipStream_t>(stream);
    int st = CAGEO_OK;
    char magic[8];
    int32_t hdr[2];
    if (!cageo::read_exact(f, magic, 8) || std::memcmp(magic, "CGSURR01", 8) != 0 ||
        !cageo::read_exact(f, hdr, sizeof(hdr)) || hdr[0] <= 0 || hdr[0] % 32 || hdr[1] <= 0 ||
        hdr[1] > 256)
        st = CAGEO_ERR_ARG;
    if (st == CAGEO_OK) s->base = hdr[0];
    for (int t = 0; st == CAGEO_OK && t < hdr[1]; ++t) {
        int32_t ln = 0, meta[2];
        if (!cageo::read_exact(f, &ln, 4) || ln <= 0 || ln > 64) {
            st = CAGEO_ERR_ARG;
            break;
        }
        std::string name(ln, '\0');
        if (!cageo::read_exact(f, &name[0], ln) || !cageo::read_exact(f, meta, 8) || meta[1] < 1 ||
            meta[1] > 4 || (meta[0] != 0 && meta[0] != 1)) {
            st = CAGEO_ERR_ARG;
            break;
        }
        std::vector<int> dims(meta[1]);
        size_t nel = 1;
        for (int d = 0; d < meta[1]; ++d) {
            int32_t v = 0;
            if (!cageo::read_exact(f, &v, 4) || v <= 0 || v > (1 << 20)) {
                st = CAGEO_ERR_ARG;
                break;
            }
            dims[d] = v;
            nel *= (size_t)v;
        }
        if (st != CAGEO_OK) break;
        const size_t bytes = nel * (meta[0] == 1 ? 2 : 4);
        std::vector<char> host(bytes);
        void* d = nullptr;
        if (!cageo::read_exact(f, host.data(), bytes)) st = CAGEO_ERR_ARG;
        else if (hipMalloc(&d, bytes) != hipSuccess) st = CAGEO_ERR_NOMEM;
        else if (hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) st = CAGEO_ERR_HIP;
        if (d) s->dev[name] = d;
        s->shape[name] = dims;
    }
    std::fclose(f);
    // every layer of the architecture, with the shapes it implies
    const int b = s->base;
    for (const auto& L : cageo::LAYERS) {
        if (st != CAGEO_OK) break;
        const std::string n(L.name);
        auto w = s->shape.find(n + ".w");
        auto bb = s->shape.find(n + ".b");
        if (w == s->shape.end() || bb == s->shape.end() || w->second.size() != 3 ||
            w->second[0] != 9 || bb->second[0] != w->second[1] || w->second[1] % 32 ||
            w->second[2] % 32)
            st = CAGEO_ERR_ARG;
    }
    if (st == CAGEO_OK && (!s->shape.count("head.w") || s->shape["head.w"][0] != b ||
                           !s->shape.count("head.b")))
        st = CAGEO_ERR_ARG;
    if (st == CAGEO_OK &&
        (hipMalloc(reinterpret_cast<void**>(&s->part), cageo::RED_BLOCKS * 3 * sizeof(double)) !=
             hipSuccess ||
         hipMalloc(reinterpret_cast<void**>(&s->stats), 2 * sizeof(float)) != hipSuccess))
        st = CAGEO_ERR_NOMEM;
    if (st != CAGEO_OK) {
        cageo_surrogate_destroy(s);
        return st;
    }
    *out = s;
    return CAGEO_OK;
}

int cageo_surrogate_forward(cageo_surrogate* s, const uint16_t* d_x, int h, int w, float* d_prob) {
    if (!s || !d_x || !d_prob || h <= 0 || w <= 0 || h % 8 || w % 8) return CAGEO_ERR_ARG;
    CAGEO_TRY(cageo::ensure_workspace(s, h, w));
    CAGEO_HIP_TRY(hipMemcpyAsync(s->act[0], d_x, (size_t)h * w * cageo::IN_PAD * sizeof(uint16_t),
                                 hipMemcpyDeviceToDevice, s->stream));
    CAGEO_TRY(cageo::forward(s));
    CAGEO_HIP_TRY(hipMemcpyAsync(d_prob, s->prob_pad, (size_t)h * w * sizeof(float),
                                 hipMemcpyDeviceToDevice, s->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(s->stream));
    return CAGEO_OK;
}

int cageo_surrogate_predict(cageo_surrogate* s, const float* d_dem, const float* d_stream, int W,
                            int H, float nodata_dem, float nodata_stream, float rmse,
                            float practical_range, float* d_prob) {
    if (!s || !d_dem || !d_stream || !d_prob || W <= 0 || H <= 0) return CAGEO_ERR_ARG;
    const int hp = (H + 7) & ~7, wp = (W + 7) & ~7;  // the network halves three times
    CAGEO_TRY(cageo::ensure_workspace(s, hp, wp));
    const long long n = (long long)W * H;
    hipLaunchKernelGGL(cageo::k_dem_partials, dim3(cageo::RED_BLOCKS), dim3(256), 0, s->stream,
                       d_dem, n, nodata_dem, s->part);
    CAGEO_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(cageo::k_dem_stats, dim3(1), dim3(64), 0, s->stream, s->part,
                       cageo::RED_BLOCKS, s->stats);
    CAGEO_HIP_TRY(hipGetLastError());
    const std::vector<int>& e1 = s->shape["enc1a.w"];
    if (s->base == 32 && e1[1] == 32 && e1[2] == cageo::IN_PAD) {
        // enc1a from the rasters (its output: the first level's temporary)
        hipLaunchKernelGGL(cageo::k_first_layer,
                           dim3((unsigned)((wp + cageo::FL_TX - 1) / cageo::FL_TX),
                                (unsigned)(hp / cageo::FL_TY)),
                           dim3(256), 0, s->stream, d_dem, d_stream, W, H, wp, hp, nodata_dem,
                           nodata_stream, rmse / 3.0f, practical_range / 200.0f, s->stats,
                           static_cast<const uint16_t*>(s->dev["enc1a.w"]),
                           static_cast<const float*>(s->dev["enc1a.b"]), s->act[1]);
        CAGEO_HIP_TRY(hipGetLastError());
        CAGEO_TRY(cageo::forward(s, true));
    } else {
        hipLaunchKernelGGL(cageo::k_make_inputs, dim3(2048), dim3(256), 0, s->stream, d_dem,
                           d_stream, W, H, wp, hp, nodata_dem, nodata_stream, rmse / 3.0f,
                           practical_range / 200.0f, s->stats, s->act[0]);
        CAGEO_HIP_TRY(hipGetLastError());
        CAGEO_TRY(cageo::forward(s));
    }
    hipLaunchKernelGGL(cageo::k_crop, dim3(2048), dim3(256), 0, s->stream, s->prob_pad, wp,
                       d_prob, W, H);
    CAGEO_HIP_TRY(hipGetLastError());
    CAGEO_HIP_TRY(hipStreamSynchronize(s->stream));
    return CAGEO_OK;
}

int cageo_surrogate_check(cageo_surrogate* s, cageo_context* ctx, const float* d_prob, int k,
                          uint64_t seed, double bias_tol, double rms_tol, float* d_scratch,
                          int* accepted, double* bias, double* rms) {
    if (!s) return CAGEO_ERR_ARG;
    return cageo_check_probability_map(ctx, d_prob, k, seed, 0, bias_tol, rms_tol, d_scratch,
                                       accepted, bias, rms);
}

}  // extern "C"
