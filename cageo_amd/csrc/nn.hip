// Kernels of the border-probability surrogate (a small U-Net): BF16 operands,
// FP32 accumulation on the MFMA units (v_mfma_f32_16x16x32_bf16), NHWC
// activations with channel counts that are multiples of 32.
//
// No reference counterpart: the reference has no learned part.  The network
// predicts the per-cell border frequency of the Monte Carlo from the DEM, the
// stream mask and the error model; the Monte Carlo generates its training
// maps and stays the judge of its output.
//
// 3x3 convolution as an implicit GEMM: rows = pixels, columns = output
// channels, K = 9 taps x input channels.  The input of a layer may be the
// concatenation of two tensors (decoder: upsampled deeper features and the
// skip connection), and the first of them may be read at half resolution
// (nearest upsampling folded into the load), so neither the concatenation
// nor the upsampled tensor is ever stored.
//
// MFMA operand layout (16x16x32, wave64): lane l holds 8 consecutive K values
// of row (l % 16) of A and of column (l % 16) of B, K block (l / 16); the
// accumulator holds rows 4*(l / 16) + r, r = 0..3, of column l % 16.
#include "common.hpp"

namespace cageo {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct ConvIn {
    const uint16_t* in1;  // [N][H1][W1][C1]; H1 = H/2, W1 = W/2 when up1
    const uint16_t* in2;  // [N][H][W][C2] or null
    int C1, C2, up1;
};

// 8 consecutive input channels [c, c+8) of pixel (n, y, x) of the (possibly
// concatenated, possibly upsampled) input; zero outside the image.
__device__ inline uint4 load_in8(const ConvIn& ci, int n, int y, int x, int H, int W, int c) {
    if (y < 0 || x < 0 || y >= H || x >= W) return make_uint4(0, 0, 0, 0);
    if (c < ci.C1) {
        const int H1 = ci.up1 ? (H + 1) / 2 : H, W1 = ci.up1 ? (W + 1) / 2 : W;
        const int yy = ci.up1 ? y / 2 : y, xx = ci.up1 ? x / 2 : x;
        return *reinterpret_cast<const uint4*>(ci.in1 + (((size_t)n * H1 + yy) * W1 + xx) * ci.C1 + c);
    }
    return *reinterpret_cast<const uint4*>(ci.in2 + (((size_t)n * H + y) * W + x) * ci.C2 + (c - ci.C1));
}

__device__ inline bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Forward 3x3 convolution (stride 1, zero padding 1).  One wave: 16*MT
// consecutive pixels x 16*NT output channels (MT x NT accumulator tiles, so
// each operand fragment loaded feeds NT or MT MFMAs); a workgroup of 4 waves
// covers 64*MT pixels.
// w: [9][Cout][Cin] bf16 (Cin = C1 + C2), bias: [Cout] fp32 or null.
// out: [N][H][W][Cout] bf16 (ReLU applied when relu), and/or outf fp32.
template <int NT, int MT>
__global__ __launch_bounds__(256) void k_conv3x3(ConvIn ci, const uint16_t* __restrict__ w,
                                                 const float* __restrict__ bias,
                                                 uint16_t* __restrict__ out,
                                                 float* __restrict__ outf, int N, int H, int W,
                                                 int Cout, int relu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long M = (long long)N * H * W;
    const long long p0 = ((long long)blockIdx.x * 4 + wave) * 16 * MT;
    const int co0 = blockIdx.y * 16 * NT;
    if (p0 >= M) return;  // whole wave out of range (uniform)
    const int Cin = ci.C1 + ci.C2;
    const int row = lane & 15, kb = lane >> 4;
    int n[MT], y[MT], x[MT];
    bool pvalid[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const long long p = p0 + 16 * m + row;
        pvalid[m] = p < M;
        n[m] = y[m] = x[m] = 0;
        if (pvalid[m]) {
            n[m] = (int)(p / ((long long)H * W));
            const int r = (int)(p - (long long)n[m] * H * W);
            y[m] = r / W;
            x[m] = r - y[m] * W;
        }
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const uint16_t* wt = w + (size_t)tap * Cout * Cin;
        for (int c = 0; c < Cin; c += 32) {
            const int cc = c + 8 * kb;
            bf16x8 a[MT], b[NT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                a[m] = as_bf16x8(pvalid[m] ? load_in8(ci, n[m], y[m] + dy, x[m] + dx, H, W, cc)
                                           : make_uint4(0, 0, 0, 0));
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int co = co0 + t * 16 + row;
                b[t] = as_bf16x8(*reinterpret_cast<const uint4*>(wt + (size_t)co * Cin + cc));
            }
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[t], acc[m][t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + t * 16 + row;  // accumulator column = lane % 16
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long long pp = p0 + 16 * m + 4 * kb + r;  // accumulator row
                if (pp >= M) continue;
                float v = acc[m][t][r] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                if (out) out[pp * Cout + co] = f2bf(v);
                if (outf) outf[pp * Cout + co] = v;
            }
    }
}

// The same convolution with its operands staged in LDS, for W % 64 == 0 and
// H % (4 RW) == 0.  A workgroup computes a tile of 4 RW rows x 64 pixels
// (wave w: rows w RW .. w RW + RW - 1, 4 accumulator tiles of 16 pixels
// each) x 16*NT output channels; with RW = 2 each weight fragment read
// feeds twice the MFMAs (the 32-channel layers).  Per 32-channel slice of
// the input it stages the (4 RW + 2) x 66 input pixels the 9 taps read (zero outside the image, upsampling and concatenation folded
// into the loads) and the slice's weights of all 9 taps, then runs the 9
// taps from LDS: each staged value is read from memory once per
// workgroup instead of once per tap and wave.  Pixel rows in LDS are
// padded to 40 bf16 (80 bytes) so 16 lanes reading 16 consecutive pixels
// hit 16 distinct 4-bank groups.
//
// Without an fp32 output the results go out through LDS: the workgroup's
// bias + ReLU outputs, rounded to bf16, are staged there, then written
// with 16-byte stores (a pixel's channels are contiguous), and the fused
// consumers of the output read them there (ConvFuse): the 2x2 average pool
// and the 1x1 head, each in the same arithmetic as its own kernel
// (k_avgpool2, k_head), so a fused layer gives the same bits as the layer
// followed by that kernel.
constexpr int CL_TX = 64;

// Place (in bf16) of 16-byte chunk c (0..3) of staged row r of the LDS
// convolution: a row is one pixel's (or one tap and output channel's) 32
// channels, 64 bytes with no padding, and its chunks are permuted by an XOR
// with bits 1-2 of the row number, so the rows a fragment read touches at
// once fall in distinct LDS banks.
__device__ inline int cl_at(int r, int c) { return r * 32 + 8 * (c ^ ((r >> 1) & 3)); }

struct ConvFuse {
    uint16_t* pool;       // [N][H/2][W/2][Cout]: k_avgpool2 of the output, or null
    const float* head_w;  // [Cout] and [1]: the 1x1 head of k_head
    const float* head_b;
    float* prob;          // [N][H][W] head output, or null (then head_w/head_b unused)
    int sigmoid;
};

template <int NT, int RW, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_conv3x3_lds(ConvIn ci, const uint16_t* __restrict__ w,
                                                     const float* __restrict__ bias,
                                                     uint16_t* __restrict__ out,
                                                     float* __restrict__ outf, ConvFuse fz,
                                                     int N, int H, int W, int Cout, int relu) {
    constexpr int TY = 4 * RW, AX = CL_TX + 2, AY = TY + 2, CO = 16 * NT;
    constexpr int TAP_UNROLL = RW == 1 ? 9 : 1;  // two rows a wave: registers run short
    constexpr int SA_N = AY * AX * 32, SB_N = 9 * CO * 32;
    constexpr int NPX = TY * CL_TX;      // pixels of the output tile
    constexpr int PS = CO + 8;           // staged pixel stride (bf16), padded by 16 bytes
    constexpr int HP = NPX + 2;          // staged channel stride of the head's layout
    static_assert(NPX * PS <= SA_N + SB_N && CO * HP <= SA_N + SB_N, "epilogue staging");
    __shared__ __align__(16) uint16_t smem[SA_N + SB_N];
    uint16_t* const sa = smem;
    uint16_t* const sb = smem + SA_N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kb = lane >> 4;
    const int tiles_x = W / CL_TX, tiles_y = H / TY;
    const int t = blockIdx.x;
    const int n = t / (tiles_x * tiles_y);
    const int ty = (t / tiles_x) % tiles_y, tx = t % tiles_x;
    const int x0 = tx * CL_TX, y0 = ty * TY;
    const int co0 = blockIdx.y * CO;
    const int Cin = ci.C1 + ci.C2;
    f32x4 acc[4 * RW][NT];
#pragma unroll
    for (int m = 0; m < 4 * RW; ++m)
#pragma unroll
        for (int q = 0; q < NT; ++q) acc[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Slice c's loads are issued before slice c - 32's MFMAs and stored to
    // LDS after them, so their latency hides behind the arithmetic.  Input
    // AY x AX pixels x 4 groups of 8 channels, weights 9 taps x CO output
    // channels x 4 groups of 8 input channels.  What does not depend on the
    // slice is worked out once: each staged input value's pixel in in1 (at
    // half resolution with up1) and in in2, -1 outside the image, and each
    // weight value's offset.  A slice lies in one of the two inputs (their
    // channel counts are multiples of 32).
    constexpr int AN = AY * AX * 4, AIT = (AN + 255) / 256;
    constexpr int BN = 9 * CO * 4, BIT = (BN + 255) / 256;
    const int H1 = ci.up1 ? (H + 1) / 2 : H, W1 = ci.up1 ? (W + 1) / 2 : W;
    int pix1[AIT], pix2[AIT], boff[BIT];
#pragma unroll
    for (int j = 0; j < AIT; ++j) {
        const int i = threadIdx.x + 256 * j;
        const int px = i >> 2, yy = y0 - 1 + px / AX, xx = x0 - 1 + px % AX;
        const bool in = i < AN && yy >= 0 && xx >= 0 && yy < H && xx < W;
        pix1[j] = in ? (n * H1 + (ci.up1 ? yy / 2 : yy)) * W1 + (ci.up1 ? xx / 2 : xx) : -1;
        pix2[j] = in ? (n * H + yy) * W + xx : -1;
    }
#pragma unroll
    for (int j = 0; j < BIT; ++j) {
        const int i = threadIdx.x + 256 * j, r = i >> 2;  // r = tap * CO + co
        const int tap = r / CO, co = r - tap * CO;
        boff[j] = i < BN ? (tap * Cout + co0 + co) * Cin + 8 * (i & 3) : -1;
    }
    uint4 av[AIT], bv[BIT];
    auto load_slice = [&](int c) {
        const bool first = c < ci.C1;
        const uint16_t* src = first ? ci.in1 : ci.in2;
        const int cs = first ? ci.C1 : ci.C2, c0 = (first ? c : c - ci.C1) + 8 * (threadIdx.x & 3);
#pragma unroll
        for (int j = 0; j < AIT; ++j) {
            const int px = first ? pix1[j] : pix2[j];
            av[j] = px >= 0 ? *reinterpret_cast<const uint4*>(src + (size_t)px * cs + c0)
                            : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int j = 0; j < BIT; ++j)
            bv[j] = boff[j] >= 0 ? *reinterpret_cast<const uint4*>(w + boff[j] + c)
                                 : make_uint4(0u, 0u, 0u, 0u);
    };
    load_slice(0);
    for (int c = 0; c < Cin; c += 32) {
        __syncthreads();  // the last slice's reads are done
#pragma unroll
        for (int j = 0; j < AIT; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (i < AN) *reinterpret_cast<uint4*>(&sa[cl_at(i >> 2, i & 3)]) = av[j];
        }
#pragma unroll
        for (int j = 0; j < BIT; ++j) {
            const int i = threadIdx.x + 256 * j;
            if (i < BN) *reinterpret_cast<uint4*>(&sb[cl_at(i >> 2, i & 3)]) = bv[j];
        }
        __syncthreads();
        if (c + 32 < Cin) load_slice(c + 32);
#pragma unroll TAP_UNROLL
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;  // halo-shifted offsets 0..2
            bf16x8 a[4 * RW], b[NT];
#pragma unroll
            for (int m = 0; m < 4 * RW; ++m)
                a[m] = as_bf16x8(*reinterpret_cast<const uint4*>(
                    &sa[cl_at((wave * RW + m / 4 + dy) * AX + 16 * (m % 4) + row + dx, kb)]));
#pragma unroll
            for (int q = 0; q < NT; ++q)
                b[q] = as_bf16x8(*reinterpret_cast<const uint4*>(
                    &sb[cl_at(tap * CO + 16 * q + row, kb)]));
#pragma unroll
            for (int m = 0; m < 4 * RW; ++m)
#pragma unroll
                for (int q = 0; q < NT; ++q)
                    acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[q], acc[m][q], 0, 0, 0);
        }
    }
    if (outf) {  // fp32 output too: straight from the accumulators
        const long long prow = ((long long)n * H + y0 + wave * RW) * W + x0;
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int co = co0 + 16 * q + row;  // accumulator column = lane % 16
            const float bias_co = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int m = 0; m < 4 * RW; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long long pp = prow + (long long)(m / 4) * W + 16 * (m % 4) + 4 * kb + r;
                    float v = acc[m][q][r] + bias_co;
                    if (relu) v = fmaxf(v, 0.0f);
                    if (out) out[pp * Cout + co] = f2bf(v);
                    outf[pp * Cout + co] = v;
                }
        }
        return;
    }
    __syncthreads();  // every wave's operand reads of the last slice are done
    // the output tile's pixel (0, 0) in the NHWC raster; the head layout is
    // channel-major (a thread then reads its pixel's channels in order, lanes
    // on consecutive pixels), the other pixel-major
    const size_t pix0 = ((size_t)n * H + y0) * W + x0;
    const bool head = fz.prob != nullptr;
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int c = 16 * q + row;  // accumulator column = lane % 16
        const float bias_co = bias ? bias[co0 + c] : 0.0f;
#pragma unroll
        for (int m = 0; m < 4 * RW; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int px = (wave * RW + m / 4) * CL_TX + 16 * (m % 4) + 4 * kb + r;
                float v = acc[m][q][r] + bias_co;
                if (relu) v = fmaxf(v, 0.0f);
                smem[head ? c * HP + px : px * PS + c] = f2bf(v);
            }
    }
    __syncthreads();
    if (head) {  // CO == Cout (the host checks): the workgroup holds every channel
        for (int px = threadIdx.x; px < NPX; px += 256) {
            float s = fz.head_b[0];
            for (int c = 0; c < CO; ++c) s += fz.head_w[c] * bf2f(smem[c * HP + px]);
            fz.prob[pix0 + (size_t)(px / CL_TX) * W + px % CL_TX] =
                fz.sigmoid ? 1.0f / (1.0f + expf(-s)) : s;
        }
        return;
    }
    constexpr int CH = CO / 8;  // 16-byte chunks of a pixel's channels
    if (out) {
        for (int i = threadIdx.x; i < NPX * CH; i += 256) {
            const int px = i / CH, part = i - px * CH;
            const size_t g = (pix0 + (size_t)(px / CL_TX) * W + px % CL_TX) * Cout + co0 + 8 * part;
            *reinterpret_cast<uint4*>(out + g) =
                *reinterpret_cast<const uint4*>(smem + px * PS + 8 * part);
        }
    }
    if (fz.pool) {
        const int Wo = W / 2;
        const size_t q0 = ((size_t)n * (H / 2) + y0 / 2) * Wo + x0 / 2;
        for (int i = threadIdx.x; i < NPX / 4 * CH; i += 256) {
            const int pp = i / CH, part = i - pp * CH;
            const int py = pp / (CL_TX / 2), px = pp - py * (CL_TX / 2);
            const uint16_t* s00 = smem + (2 * py * CL_TX + 2 * px) * PS + 8 * part;
            const uint4 a = *reinterpret_cast<const uint4*>(s00);                 // (2y, 2x)
            const uint4 b = *reinterpret_cast<const uint4*>(s00 + PS);            // (2y, 2x+1)
            const uint4 c = *reinterpret_cast<const uint4*>(s00 + CL_TX * PS);    // (2y+1, 2x)
            const uint4 d = *reinterpret_cast<const uint4*>(s00 + (CL_TX + 1) * PS);
            const uint32_t ua[4] = {a.x, a.y, a.z, a.w}, ub[4] = {b.x, b.y, b.z, b.w};
            const uint32_t uc[4] = {c.x, c.y, c.z, c.w}, ud[4] = {d.x, d.y, d.z, d.w};
            uint32_t res[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t h2 = 0;
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
                    const int sh = 16 * hh;
                    const float sum = bf2f((uint16_t)(ua[k] >> sh)) + bf2f((uint16_t)(ub[k] >> sh)) +
                                      bf2f((uint16_t)(uc[k] >> sh)) + bf2f((uint16_t)(ud[k] >> sh));
                    h2 |= (uint32_t)f2bf(0.25f * sum) << sh;
                }
                res[k] = h2;
            }
            *reinterpret_cast<uint4*>(fz.pool + (q0 + (size_t)py * Wo + px) * Cout + co0 + 8 * part) =
                make_uint4(res[0], res[1], res[2], res[3]);
        }
    }
}

// Weight gradient of the 3x3 convolution:
//   dW[tap][co][ci] += sum_p g[p][co] * x[p shifted by tap][ci]
// (GEMM rows co, columns ci, K = pixels).  A workgroup takes one tap, a
// 32x32 (co, ci) block and a run of pixel chunks; each chunk of 32 pixels is
// staged through LDS (rows of g and of the shifted input are contiguous in
// NHWC) and read back transposed as MFMA operands.  Partial sums are added
// to the fp32 gradient with atomics.
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(ConvIn ci, const uint16_t* __restrict__ g,
                                                       float* __restrict__ dw, int N, int H, int W,
                                                       int Cout, int chunks_per_block) {
    __shared__ uint16_t sg[32][32 + 8];  // [pixel][co]
    __shared__ uint16_t sx[32][32 + 8];  // [pixel][ci]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Cin = ci.C1 + ci.C2;
    const int tap = blockIdx.z % 9;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int co0 = blockIdx.y * 32;
    const int ci0 = (blockIdx.z / 9) * 32;
    const long long M = (long long)N * H * W;
    // wave (wr, wc) computes the 16x16 block (co0 + 16 wr, ci0 + 16 wc)
    const int wr = wave >> 1, wc = wave & 1;
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    const long long chunk0 = (long long)blockIdx.x * chunks_per_block;
    for (int ch = 0; ch < chunks_per_block; ++ch) {
        const long long pbase = (chunk0 + ch) * 32;
        if (pbase >= M) break;  // uniform
        // stage: 256 threads load 32 pixels x 32 channels of g and of x
        {
            const int pl = threadIdx.x >> 3, c8 = (threadIdx.x & 7) * 4;  // 4 channels each
            const long long p = pbase + pl;
            uint2 gv = make_uint2(0, 0), xv = make_uint2(0, 0);
            if (p < M) {
                gv = *reinterpret_cast<const uint2*>(g + p * Cout + co0 + c8);
                const int n = (int)(p / ((long long)H * W));
                const int r = (int)(p - (long long)n * H * W);
                const int y = r / W, x = r - (r / W) * W;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && xx >= 0 && yy < H && xx < W) {
                    const int c = ci0 + c8;
                    if (c < ci.C1) {
                        const int H1 = ci.up1 ? (H + 1) / 2 : H, W1 = ci.up1 ? (W + 1) / 2 : W;
                        const int y1 = ci.up1 ? yy / 2 : yy, x1 = ci.up1 ? xx / 2 : xx;
                        xv = *reinterpret_cast<const uint2*>(
                            ci.in1 + (((size_t)n * H1 + y1) * W1 + x1) * ci.C1 + c);
                    } else {
                        xv = *reinterpret_cast<const uint2*>(
                            ci.in2 + (((size_t)n * H + yy) * W + xx) * ci.C2 + (c - ci.C1));
                    }
                }
            }
            *reinterpret_cast<uint2*>(&sg[pl][c8]) = gv;
            *reinterpret_cast<uint2*>(&sx[pl][c8]) = xv;
        }
        __syncthreads();
        // A[i = co][k = pixel], B[k = pixel][j = ci]; lane: row/col l % 16,
        // pixels 8*(l / 16) .. +7
        const int i = lane & 15, kb = lane >> 4;
        uint16_t av[8], bv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            av[e] = sg[8 * kb + e][16 * wr + i];
            bv[e] = sx[8 * kb + e][16 * wc + i];
        }
        uint4 a4, b4;
        a4.x = av[0] | ((uint32_t)av[1] << 16); a4.y = av[2] | ((uint32_t)av[3] << 16);
        a4.z = av[4] | ((uint32_t)av[5] << 16); a4.w = av[6] | ((uint32_t)av[7] << 16);
        b4.x = bv[0] | ((uint32_t)bv[1] << 16); b4.y = bv[2] | ((uint32_t)bv[3] << 16);
        b4.z = bv[4] | ((uint32_t)bv[5] << 16); b4.w = bv[6] | ((uint32_t)bv[7] << 16);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a4), as_bf16x8(b4), acc, 0, 0, 0);
        __syncthreads();
    }
    const int j = lane & 15, kb = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int co = co0 + 16 * wr + 4 * kb + r;
        const int cix = ci0 + 16 * wc + j;
        atomicAdd(&dw[((size_t)tap * Cout + co) * Cin + cix], acc[r]);
    }
}

// ---- elementwise / small kernels ------------------------------------------

#define GRID_STRIDE(i, n)                                                               \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)(n); \
         i += (long long)gridDim.x * blockDim.x)

// 2x2 average pooling, NHWC bf16 (H, W even).
__global__ void k_avgpool2(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int N,
                           int H, int W, int C) {
    const int Ho = H / 2, Wo = W / 2;
    GRID_STRIDE(i, (long long)N * Ho * Wo * C) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % Wo);
        p /= Wo;
        const int y = (int)(p % Ho);
        const int n = (int)(p / Ho);
        const size_t b = (((size_t)n * H + 2 * y) * W + 2 * x) * C + c;
        const float s = bf2f(in[b]) + bf2f(in[b + C]) + bf2f(in[b + (size_t)W * C]) +
                        bf2f(in[b + (size_t)W * C + C]);
        out[i] = f2bf(0.25f * s);
    }
}

// Gradient of the 2x2 average pool, added to `dx` (fp32, full resolution).
__global__ void k_avgpool2_back(const float* __restrict__ dy, float* __restrict__ dx, int N, int H,
                                int W, int C) {
    GRID_STRIDE(i, (long long)N * H * W * C) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % W);
        p /= W;
        const int y = (int)(p % H);
        const int n = (int)(p / H);
        dx[i] += 0.25f * dy[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
}

// Gradient of nearest 2x upsampling: sum of the 2x2 block (into `dx`, +=).
__global__ void k_upsample2_back(const float* __restrict__ dy, float* __restrict__ dx, int N,
                                 int H, int W, int C, int Ctot, int coff) {
    // dy: [N][H][W][Ctot] (channels coff..coff+C belong to the upsampled input)
    const int Hh = (H + 1) / 2, Wh = (W + 1) / 2;
    GRID_STRIDE(i, (long long)N * Hh * Wh * C) {
        const int c = (int)(i % C);
        long long p = i / C;
        const int x = (int)(p % Wh);
        p /= Wh;
        const int y = (int)(p % Hh);
        const int n = (int)(p / Hh);
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const int yy = 2 * y + a, xx = 2 * x + b;
                if (yy < H && xx < W) s += dy[(((size_t)n * H + yy) * W + xx) * Ctot + coff + c];
            }
        dx[i] += s;
    }
}

// dx[p][c] (+)= dy[p][coff + c] for c < C (a slice of a concatenated gradient).
__global__ void k_slice_add(const float* __restrict__ dy, float* __restrict__ dx, long long P,
                            int C, int Ctot, int coff, int accumulate) {
    GRID_STRIDE(i, P * C) {
        const long long p = i / C;
        const int c = (int)(i % C);
        const float v = dy[p * Ctot + coff + c];
        dx[i] = accumulate ? dx[i] + v : v;
    }
}

// ReLU backward into bf16 for the next dgrad: g = dy * (act > 0).
__global__ void k_relu_back_bf16(const float* __restrict__ dy, const uint16_t* __restrict__ act,
                                 uint16_t* __restrict__ g, long long n) {
    GRID_STRIDE(i, n) {
        const float v = (act[i] & 0x7fffu) && !(act[i] & 0x8000u) ? dy[i] : 0.0f;
        g[i] = f2bf(v);
    }
}

// ReLU backward and bias gradient in one pass: g = bf16(dy * (act > 0)) and
// db[c] += sum_p g[p][c] (of the rounded g).  C divides the workgroup size
// and the grid stride is a multiple of it, so a thread only ever sees
// channel threadIdx.x % C; a workgroup adds its threads' sums in LDS and
// makes one atomic per channel.
__global__ __launch_bounds__(256) void k_relu_back_bias(const float* __restrict__ dy,
                                                        const uint16_t* __restrict__ act,
                                                        uint16_t* __restrict__ g, long long n,
                                                        int C, float* __restrict__ db) {
    __shared__ float part[256];
    float s = 0.0f;
    GRID_STRIDE(i, n) {
        const float v = (act[i] & 0x7fffu) && !(act[i] & 0x8000u) ? dy[i] : 0.0f;
        const uint16_t b = f2bf(v);
        g[i] = b;
        s += bf2f(b);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.0f;
        for (int k = threadIdx.x; k < (int)blockDim.x; k += C) t += part[k];
        atomicAdd(&db[threadIdx.x], t);
    }
}

// Bias gradient: db[c] += sum_p g[p][c] (g bf16).
__global__ void k_bias_grad(const uint16_t* __restrict__ g, float* __restrict__ db, long long P,
                            int C) {
    // one block per 64-channel group, threads stride over pixels
    const int c = blockIdx.y * 64 + (threadIdx.x & 63);
    const int lanes = blockDim.x / 64;
    float s = 0.0f;
    if (c < C)
        for (long long p = (long long)blockIdx.x * lanes + (threadIdx.x >> 6); p < P;
             p += (long long)gridDim.x * lanes)
            s += bf2f(g[p * C + c]);
    if (c < C) atomicAdd(&db[c], s);
}

// Head: logit[p] = b + sum_c w[c] act[p][c] (1x1 convolution to one
// channel); with `sigmoid` it writes the probability instead.
__global__ void k_head(const uint16_t* __restrict__ act, const float* __restrict__ w,
                       const float* __restrict__ b, float* __restrict__ logit, long long P, int C,
                       int sigmoid) {
    GRID_STRIDE(p, P) {
        float s = b[0];
        // C is a multiple of 32: the pixel's channels as 16-byte loads, summed
        // in channel order
        const uint4* a = reinterpret_cast<const uint4*>(act + p * C);
        for (int c8 = 0; c8 < C / 8; ++c8) {
            const uint4 v = a[c8];
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                s += w[8 * c8 + 2 * k] * bf2f((uint16_t)(u[k] & 0xffffu));
                s += w[8 * c8 + 2 * k + 1] * bf2f((uint16_t)(u[k] >> 16));
            }
        }
        logit[p] = sigmoid ? 1.0f / (1.0f + expf(-s)) : s;
    }
}

// Loss and head gradients: q = sigmoid(logit); binary cross-entropy against
// the target frequency t; dlogit = (q - t) / P.  Writes q, the head weight and
// bias gradients and d(act) in fp32.  loss[0] += sum BCE / P.  C is a
// multiple of 32 up to HEAD_MAX_C: each thread sums its pixels' terms of
// the weight gradient in registers, a workgroup adds them up in LDS and
// makes one atomic per channel.
constexpr int HEAD_MAX_C = 128;
template <int C>
__global__ __launch_bounds__(256) void k_head_back(const uint16_t* __restrict__ act,
                                                   const float* __restrict__ w,
                                                   const float* __restrict__ logit,
                                                   const float* __restrict__ target,
                                                   float* __restrict__ q, float* __restrict__ dact,
                                                   float* __restrict__ dw, float* __restrict__ db,
                                                   float* __restrict__ loss, long long P) {
    __shared__ float part[C];
    __shared__ float red[2];
    for (int c = threadIdx.x; c < C; c += blockDim.x) part[c] = 0.0f;
    if (threadIdx.x < 2) red[threadIdx.x] = 0.0f;
    __syncthreads();
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0f;
    float lsum = 0.0f, bsum = 0.0f;
    GRID_STRIDE(p, P) {
        const float z = logit[p];
        const float qq = 1.0f / (1.0f + expf(-z));
        q[p] = qq;
        const float t = target ? target[p] : 0.0f;
        // BCE with logits, stable form: max(z,0) - z t + log(1 + e^-|z|)
        lsum += fmaxf(z, 0.0f) - z * t + log1pf(expf(-fabsf(z)));
        const float g = (qq - t) / (float)P;
        bsum += g;
        const uint4* a = reinterpret_cast<const uint4*>(act + p * C);
        float4* d = reinterpret_cast<float4*>(dact + p * C);
#pragma unroll
        for (int c8 = 0; c8 < C / 8; ++c8) {
            const uint4 v = a[c8];
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int c = 8 * c8 + 2 * k;
                acc[c] += g * bf2f((uint16_t)(u[k] & 0xffffu));
                acc[c + 1] += g * bf2f((uint16_t)(u[k] >> 16));
            }
            d[2 * c8] = make_float4(g * w[8 * c8], g * w[8 * c8 + 1], g * w[8 * c8 + 2], g * w[8 * c8 + 3]);
            d[2 * c8 + 1] = make_float4(g * w[8 * c8 + 4], g * w[8 * c8 + 5], g * w[8 * c8 + 6],
                                        g * w[8 * c8 + 7]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        float v = acc[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) atomicAdd(&part[c], v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lsum += __shfl_xor(lsum, off);
        bsum += __shfl_xor(bsum, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&red[0], lsum);
        atomicAdd(&red[1], bsum);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&dw[c], part[c]);
    if (threadIdx.x == 0) {
        atomicAdd(loss, red[0] / (float)P);
        atomicAdd(db, red[1]);
    }
}

// Adam step on fp32 master weights, refreshing the bf16 copy (may be null).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pb, long long n, float lr,
                       float b1, float b2, float eps, float bc1, float bc2) {
    GRID_STRIDE(i, n) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        const float np = p[i] - upd;
        p[i] = np;
        if (pb) pb[i] = f2bf(np);
    }
}

// fp32 -> bf16 (weights, inputs).
__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long long n) {
    GRID_STRIDE(i, n) y[i] = f2bf(x[i]);
}

// Weights of the data-gradient convolution: wt[8 - tap][ci][co] = w[tap][co][ci].
__global__ void k_flip_transpose(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                                 int Cout, int Cin) {
    GRID_STRIDE(i, 9LL * Cout * Cin) {
        const int c_in = (int)(i % Cin);
        long long r = i / Cin;
        const int co = (int)(r % Cout);
        const int tap = (int)(r / Cout);
        wt[((size_t)(8 - tap) * Cin + c_in) * Cout + co] = w[i];
    }
}

int grid_for(long long n) {
    long long g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

}  // namespace

// The LDS-staged convolution applies: whole 64-pixel row segments, whole
// 4-row blocks, output channels in blocks of 32.
inline bool lds_conv_applies(int h, int wdt, int cout) {
    return wdt % CL_TX == 0 && h % 4 == 0 && cout % 32 == 0;
}
// Output channels of one workgroup of the LDS-staged convolution.
inline int lds_conv_block(int cout) { return cout % 64 == 0 ? 64 : 32; }

// Workgroups of 4 x 64 pixels x 64 output channels (two a CU: 210 VGPRs),
// or of 4 x 64 pixels x 32 output channels, three a CU (146 VGPRs, 43.8 KB of
// LDS): a 32-channel layer has few MFMAs per staged byte, so its workgroups
// mostly wait on their loads and stores, and a third one a CU hides more of
// that than taller tiles would save (the three such layers of a 4096^2
// prediction: 891 -> 843 us a call against 8 x 64-pixel tiles at two a CU).
int launch_conv_lds(void* stream, const ConvIn& ci, const uint16_t* w, const float* bias,
                    uint16_t* out, float* outf, const ConvFuse& fz, int n, int h, int wdt, int cout,
                    int relu) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const unsigned tx = (unsigned)(wdt / CL_TX);
    if (cout % 64 == 0)
        hipLaunchKernelGGL((k_conv3x3_lds<4, 1, 2>), dim3((unsigned)n * (h / 4) * tx, cout / 64),
                           dim3(256), 0, st, ci, w, bias, out, outf, fz, n, h, wdt, cout, relu);
    else
        hipLaunchKernelGGL((k_conv3x3_lds<2, 1, 3>), dim3((unsigned)n * (h / 4) * tx, cout / 32),
                           dim3(256), 0, st, ci, w, bias, out, outf, fz, n, h, wdt, cout, relu);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // namespace cageo

// ---- C ABI ------------------------------------------------------------------
using namespace cageo;

extern "C" {

int cageo_nn_conv3x3(void* stream, const uint16_t* in1, int c1, int up1, const uint16_t* in2,
                     int c2, const uint16_t* w, const float* bias, uint16_t* out, float* outf,
                     int n, int h, int wdt, int cout, int relu) {
    if (!in1 || !w || (!out && !outf) || n <= 0 || h <= 0 || wdt <= 0) return CAGEO_ERR_ARG;
    if (c1 <= 0 || c1 % 32 || c2 < 0 || c2 % 32 || (c2 && !in2) || cout % 16) return CAGEO_ERR_ARG;
    ConvIn ci{in1, in2, c1, c2, up1};
    const long long M = (long long)n * h * wdt;
    if (lds_conv_applies(h, wdt, cout))
        return launch_conv_lds(stream, ci, w, bias, out, outf, ConvFuse{}, n, h, wdt, cout, relu);
    // output channel tile: 64 (NT 4), 32 (NT 2) or 16 (NT 1); 4 pixel tiles
    // per wave (256 pixels per workgroup; 8 lowered occupancy and time)
    constexpr int MT = 4;
    const int nt = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
    dim3 grid((unsigned)((M + 64 * MT - 1) / (64 * MT)), (unsigned)(cout / (16 * nt)));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (nt == 4)
        hipLaunchKernelGGL((k_conv3x3<4, MT>), grid, dim3(256), 0, s, ci, w, bias, out, outf, n, h,
                           wdt, cout, relu);
    else if (nt == 2)
        hipLaunchKernelGGL((k_conv3x3<2, MT>), grid, dim3(256), 0, s, ci, w, bias, out, outf, n, h,
                           wdt, cout, relu);
    else
        hipLaunchKernelGGL((k_conv3x3<1, MT>), grid, dim3(256), 0, s, ci, w, bias, out, outf, n, h,
                           wdt, cout, relu);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_conv3x3_fusable(int h, int wdt, int cout, int head) {
    if (h <= 0 || wdt <= 0 || cout <= 0 || !lds_conv_applies(h, wdt, cout)) return 0;
    return head ? (lds_conv_block(cout) == cout ? 1 : 0) : 1;
}

int cageo_nn_conv3x3_fused(void* stream, const uint16_t* in1, int c1, int up1,
                           const uint16_t* in2, int c2, const uint16_t* w, const float* bias,
                           uint16_t* out, uint16_t* pool, const float* head_w,
                           const float* head_b, float* prob, int head_sigmoid, int n, int h,
                           int wdt, int cout, int relu) {
    if (!in1 || !w || n <= 0 || h <= 0 || wdt <= 0) return CAGEO_ERR_ARG;
    if (c1 <= 0 || c1 % 32 || c2 < 0 || c2 % 32 || (c2 && !in2)) return CAGEO_ERR_ARG;
    if (!out && !pool && !prob) return CAGEO_ERR_ARG;
    if (prob && (out || pool || !head_w || !head_b)) return CAGEO_ERR_ARG;
    if (!cageo_nn_conv3x3_fusable(h, wdt, cout, prob ? 1 : 0)) return CAGEO_ERR_ARG;
    ConvIn ci{in1, in2, c1, c2, up1};
    ConvFuse fz{pool, head_w, head_b, prob, head_sigmoid};
    return launch_conv_lds(stream, ci, w, bias, out, nullptr, fz, n, h, wdt, cout, relu);
}

int cageo_nn_conv3x3_wgrad(void* stream, const uint16_t* in1, int c1, int up1,
                           const uint16_t* in2, int c2, const uint16_t* g, float* dw, int n, int h,
                           int wdt, int cout) {
    if (!in1 || !g || !dw || c1 <= 0 || c1 % 32 || c2 % 32 || cout % 32) return CAGEO_ERR_ARG;
    ConvIn ci{in1, in2, c1, c2, up1};
    const long long M = (long long)n * h * wdt;
    const long long chunks = (M + 31) / 32;
    const int cin = c1 + c2;
    // enough workgroups to fill the chip, each over a run of pixel chunks
    const long long blocks_xy = (long long)(cout / 32) * 9 * (cin / 32);
    long long bx = (2048 + blocks_xy - 1) / blocks_xy;
    if (bx > chunks) bx = chunks;
    if (bx < 1) bx = 1;
    const int per = (int)((chunks + bx - 1) / bx);
    dim3 grid((unsigned)((chunks + per - 1) / per), (unsigned)(cout / 32), (unsigned)(9 * (cin / 32)));
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       ci, g, dw, n, h, wdt, cout, per);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_flip_transpose(void* stream, const uint16_t* w, uint16_t* wt, int cout, int cin) {
    hipLaunchKernelGGL(k_flip_transpose, dim3(grid_for(9LL * cout * cin)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), w, wt, cout, cin);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_avgpool2(void* stream, const uint16_t* in, uint16_t* out, int n, int h, int w, int c) {
    if (h % 2 || w % 2) return CAGEO_ERR_ARG;
    hipLaunchKernelGGL(k_avgpool2, dim3(grid_for((long long)n * h * w * c / 4)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), in, out, n, h, w, c);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_avgpool2_back(void* stream, const float* dy, float* dx, int n, int h, int w, int c) {
    hipLaunchKernelGGL(k_avgpool2_back, dim3(grid_for((long long)n * h * w * c)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dy, dx, n, h, w, c);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_upsample2_back(void* stream, const float* dy, float* dx, int n, int h, int w, int c,
                            int ctot, int coff) {
    hipLaunchKernelGGL(k_upsample2_back, dim3(grid_for((long long)n * ((h + 1) / 2) * ((w + 1) / 2) * c)),
                       dim3(256), 0, reinterpret_cast<hipStream_t>(stream), dy, dx, n, h, w, c,
                       ctot, coff);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_slice_add(void* stream, const float* dy, float* dx, long long p, int c, int ctot,
                       int coff, int accumulate) {
    hipLaunchKernelGGL(k_slice_add, dim3(grid_for(p * c)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dy, dx, p, c, ctot, coff, accumulate);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_relu_back(void* stream, const float* dy, const uint16_t* act, uint16_t* g,
                       long long n) {
    hipLaunchKernelGGL(k_relu_back_bf16, dim3(grid_for(n)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dy, act, g, n);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_bias_grad(void* stream, const uint16_t* g, float* db, long long p, int c) {
    dim3 grid((unsigned)std::min<long long>(1024, (p + 3) / 4), (unsigned)((c + 63) / 64));
    hipLaunchKernelGGL(k_bias_grad, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), g,
                       db, p, c);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_head(void* stream, const uint16_t* act, const float* w, const float* b, float* logit,
                  long long p, int c, int sigmoid) {
    if (!act || !w || !b || !logit || p < 0 || c <= 0 || c % 32) return CAGEO_ERR_ARG;
    hipLaunchKernelGGL(k_head, dim3(grid_for(p)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), act, w, b, logit, p, c, sigmoid);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_head_back(void* stream, const uint16_t* act, const float* w, const float* logit,
                       const float* target, float* q, float* dact, float* dw, float* db,
                       float* loss, long long p, int c) {
    if (!act || !w || !logit || !q || !dact || !dw || !db || !loss || p <= 0) return CAGEO_ERR_ARG;
    const int grid = (int)std::min<long long>(1024, (p + 255) / 256);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (c == 32)
        hipLaunchKernelGGL(k_head_back<32>, dim3(grid), dim3(256), 0, s, act, w, logit, target, q,
                           dact, dw, db, loss, p);
    else if (c == 64)
        hipLaunchKernelGGL(k_head_back<64>, dim3(grid), dim3(256), 0, s, act, w, logit, target, q,
                           dact, dw, db, loss, p);
    else if (c == 96)
        hipLaunchKernelGGL(k_head_back<96>, dim3(grid), dim3(256), 0, s, act, w, logit, target, q,
                           dact, dw, db, loss, p);
    else if (c == HEAD_MAX_C)
        hipLaunchKernelGGL(k_head_back<HEAD_MAX_C>, dim3(grid), dim3(256), 0, s, act, w, logit,
                           target, q, dact, dw, db, loss, p);
    else
        return CAGEO_ERR_ARG;
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_relu_back_bias(void* stream, const float* dy, const uint16_t* act, uint16_t* g,
                            long long p, int c, float* db) {
    if (!dy || !act || !g || !db || p <= 0 || c <= 0 || 256 % c) return CAGEO_ERR_ARG;
    const long long n = p * c;
    const long long blocks = std::min<long long>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(k_relu_back_bias, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dy, act, g, n, c, db);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_adam(void* stream, float* p, const float* g, float* m, float* v, uint16_t* pb,
                  long long n, float lr, float b1, float b2, float eps, int step) {
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), p, g, m, v, pb, n, lr, b1, b2, eps,
                       bc1, bc2);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int cageo_nn_to_bf16(void* stream, const float* x, uint16_t* y, long long n) {
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), x, y, n);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

}  // extern "C"
