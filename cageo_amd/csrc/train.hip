// Training of the border-probability surrogate behind the C ABI: forward,
// backward and Adam of the U-Net on the library's BF16 MFMA kernels (nn.hip),
// and a seeded training loop over a set of Monte Carlo-labelled tiles.
//
// No reference counterpart: the reference has no learned part.  The
// network is the one cageo_surrogate_create loads (surrogate.hip): base b
// channels, enc1 32->b->b, enc2 b->2b->2b, enc3 2b->4b->4b, mid 4b->8b->8b,
// dec3 [8b up | 4b]->4b->4b, dec2 [4b up | 2b]->2b->2b, dec1 [2b up | b]->b->b,
// head 1x1 b->1.  Master weights, gradients and Adam moments are fp32; the
// convolutions read bf16 copies of the weights and bf16 activations with
// fp32 accumulation.  The loss is the binary cross-entropy of the sigmoid
// output against the Monte Carlo frequency.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"

namespace cageo {

namespace {

constexpr int TR_IN_PAD = 32;  // input channels the first convolution reads
constexpr int TR_IN_CH = 4;    // of which used: DEM, stream mask, rmse / 3, range / 200

struct TrLayer {
    const char* name;
    int c1m, up1, c2m, com, level;  // channels as multiples of b (c1m = 0: the 32 input planes)
};
// in execution order
constexpr TrLayer TR_LAYERS[] = {
    {"enc1a", 0, 0, 0, 1, 0}, {"enc1b", 1, 0, 0, 1, 0}, {"enc2a", 1, 0, 0, 2, 1},
    {"enc2b", 2, 0, 0, 2, 1}, {"enc3a", 2, 0, 0, 4, 2}, {"enc3b", 4, 0, 0, 4, 2},
    {"mida", 4, 0, 0, 8, 3},  {"midb", 8, 0, 0, 8, 3},  {"dec3a", 8, 1, 4, 4, 2},
    {"dec3b", 4, 0, 0, 4, 2}, {"dec2a", 4, 1, 2, 2, 1}, {"dec2b", 2, 0, 0, 2, 1},
    {"dec1a", 2, 1, 1, 1, 0}, {"dec1b", 1, 0, 0, 1, 0}};
constexpr int TR_NLAYERS = sizeof(TR_LAYERS) / sizeof(TR_LAYERS[0]);

struct Param {
    std::string name;
    std::vector<int> shape;
    long long n = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* pb = nullptr;  // bf16 copy (convolution weights only)
};

__device__ inline uint16_t tr_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

#define TR_STRIDE(i, n)                                                                    \
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)(n); \
         i += (long long)gridDim.x * blockDim.x)

int tr_grid(long long n) {
    const long long g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 16384 ? 16384 : g));
}

__global__ void k_tr_zero(float* __restrict__ p, long long n) { TR_STRIDE(i, n) p[i] = 0.0f; }

// The padded input channels of the first layer are not trained: their
// gradient (rows of `cin` channels, from channel `from` on) is cleared.
__global__ void k_tr_zero_pad(float* __restrict__ g, long long rows, int cin, int from) {
    TR_STRIDE(i, rows * (cin - from)) {
        const long long r = i / (cin - from);
        g[r * cin + from + (int)(i % (cin - from))] = 0.0f;
    }
}

// Batch assembly: sample i of the batch is tile sel[2i] of the set (inputs
// x4 [tiles][h][w][4] fp32, targets y [tiles][h][w]) under dihedral
// transform sel[2i+1] (bits 0-1 quarter turns, bit 2 mirror; quarter turns
// only when h == w).  Writes the bf16 input planes padded to 32 channels
// and the fp32 target.
__global__ void k_tr_gather(const float* __restrict__ x4, const float* __restrict__ y,
                            const int* __restrict__ sel, int n, int h, int w,
                            uint16_t* __restrict__ xb, float* __restrict__ yb) {
    TR_STRIDE(i, (long long)n * h * w) {
        const int s = (int)(i / ((long long)h * w));
        const int r = (int)(i - (long long)s * h * w);
        const int oy = r / w, ox = r - oy * w;
        const int tile = sel[2 * s], tf = sel[2 * s + 1];
        int yy = oy, xx = (tf & 4) ? w - 1 - ox : ox;
        int sy = yy, sx = xx;
        switch (tf & 3) {
            case 1: sy = xx; sx = w - 1 - yy; break;
            case 2: sy = h - 1 - yy; sx = w - 1 - xx; break;
            case 3: sy = h - 1 - xx; sx = yy; break;
            default: break;
        }
        const size_t src = ((size_t)tile * h + sy) * w + sx;
        const float4 v = *reinterpret_cast<const float4*>(x4 + 4 * src);
        uint4* o = reinterpret_cast<uint4*>(xb + (size_t)i * TR_IN_PAD);
        o[0] = make_uint4((uint32_t)tr_bf16(v.x) | ((uint32_t)tr_bf16(v.y) << 16),
                          (uint32_t)tr_bf16(v.z) | ((uint32_t)tr_bf16(v.w) << 16), 0u, 0u);
        o[1] = make_uint4(0u, 0u, 0u, 0u);
        o[2] = make_uint4(0u, 0u, 0u, 0u);
        o[3] = make_uint4(0u, 0u, 0u, 0u);
        yb[i] = y ? y[src] : 0.0f;
    }
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

bool tr_read(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

}  // namespace

}  // namespace cageo

struct cageo_trainer {
    hipStream_t stream = nullptr;
    int base = 0;
    int step = 0;  // Adam steps taken
    std::vector<cageo::Param> params;
    // workspace for batches of n tiles of h x w
    int n = 0, h = 0, w = 0;
    std::vector<void*> bufs;
    uint16_t* xb = nullptr;  // input planes, bf16 [n][h][w][32]
    float* yb = nullptr;     // targets
    std::vector<uint16_t*> act;  // forward activations, see forward()
    float *logit = nullptr, *q = nullptr;
    std::vector<float*> gr;      // backward gradients, see backward()
    uint16_t* gtmp = nullptr;    // bf16 output gradient of the layer being differentiated
    uint16_t* wt = nullptr;      // flipped, transposed weights of that layer
    float* loss = nullptr;       // per-step losses of the running loop
    int loss_cap = 0;
    int* sel = nullptr;          // batch selections of the running loop
    long long sel_cap = 0;
};

namespace cageo {

namespace {

Param* find(cageo_trainer* t, const std::string& name) {
    for (Param& p : t->params)
        if (p.name == name) return &p;
    return nullptr;
}

void free_batch(cageo_trainer* t) {
    for (void* p : t->bufs) (void)hipFree(p);
    t->bufs.clear();
    t->act.clear();
    t->gr.clear();
    t->xb = nullptr;
    t->yb = nullptr;
    t->logit = t->q = nullptr;
    t->gtmp = t->wt = nullptr;
    t->n = t->h = t->w = 0;
}

template <class T>
int alloc(cageo_trainer* t, T** out, size_t elems) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(elems, 1) * sizeof(T)) != hipSuccess) return CAGEO_ERR_NOMEM;
    t->bufs.push_back(p);
    *out = static_cast<T*>(p);
    return CAGEO_OK;
}

// Activation and gradient buffers for batches of n x h x w (h, w multiples
// of 8).
int ensure_batch(cageo_trainer* t, int n, int h, int w) {
    if (t->n == n && t->h == h && t->w == w) return CAGEO_OK;
    free_batch(t);
    const int b = t->base;
    const size_t P = (size_t)n * h * w;
    auto px = [&](int lv) { return P >> (2 * lv); };
    CAGEO_TRY(alloc(t, &t->xb, P * TR_IN_PAD));
    CAGEO_TRY(alloc(t, &t->yb, P));
    // forward: one bf16 output per layer (execution order) and the three
    // pooled tensors
    for (const TrLayer& L : TR_LAYERS) {
        uint16_t* a = nullptr;
        CAGEO_TRY(alloc(t, &a, px(L.level) * L.com * b));
        t->act.push_back(a);
    }
    for (int lv = 1; lv <= 3; ++lv) {  // p1, p2, p3: the pooled encoder outputs
        uint16_t* a = nullptr;
        CAGEO_TRY(alloc(t, &a, px(lv) * (b << (lv - 1))));
        t->act.push_back(a);
    }
    CAGEO_TRY(alloc(t, &t->logit, P));
    CAGEO_TRY(alloc(t, &t->q, P));
    // backward: per layer the gradient of its output (fp32) and, for the
    // decoder's first layers, of its concatenated input
    for (const TrLayer& L : TR_LAYERS) {
        float* g = nullptr;
        CAGEO_TRY(alloc(t, &g, px(L.level) * L.com * b));
        t->gr.push_back(g);
    }
    for (const TrLayer& L : TR_LAYERS) {
        float* g = nullptr;
        const int cin = L.c1m == 0 ? TR_IN_PAD : (L.c1m + L.c2m) * b;
        CAGEO_TRY(alloc(t, &g, L.up1 ? px(L.level) * cin : 1));
        t->gr.push_back(g);
    }
    for (int lv = 1; lv <= 3; ++lv) {  // gradients of the pooled tensors
        float* g = nullptr;
        CAGEO_TRY(alloc(t, &g, px(lv) * (b << (lv - 1))));
        t->gr.push_back(g);
    }
    CAGEO_TRY(alloc(t, &t->gtmp, P * b));  // the largest layer output: b channels at full size
    size_t wmax = 0;  // the largest 9 x cin x cout
    for (const TrLayer& L : TR_LAYERS)
        wmax = std::max(wmax, (size_t)9 * (L.c1m == 0 ? TR_IN_PAD : (L.c1m + L.c2m) * b) * L.com * b);
    CAGEO_TRY(alloc(t, &t->wt, wmax));
    t->n = n;
    t->h = h;
    t->w = w;
    return CAGEO_OK;
}

enum { E1A, E1B, E2A, E2B, E3A, E3B, MA, MB, D3A, D3B, D2A, D2B, D1A, D1B };

// Inputs of layer i: in1 (c1 channels; upsampled with up1) and in2 (c2).
struct LayerIn {
    const uint16_t* in1;
    int c1, up1;
    const uint16_t* in2;
    int c2;
};
LayerIn layer_in(cageo_trainer* t, int i) {
    const int b = t->base;
    const auto& A = t->act;
    const uint16_t* P1 = A[TR_NLAYERS], *P2 = A[TR_NLAYERS + 1], *P3 = A[TR_NLAYERS + 2];
    switch (i) {
        case E1A: return {t->xb, TR_IN_PAD, 0, nullptr, 0};
        case E1B: return {A[E1A], b, 0, nullptr, 0};
        case E2A: return {P1, b, 0, nullptr, 0};
        case E2B: return {A[E2A], 2 * b, 0, nullptr, 0};
        case E3A: return {P2, 2 * b, 0, nullptr, 0};
        case E3B: return {A[E3A], 4 * b, 0, nullptr, 0};
        case MA: return {P3, 4 * b, 0, nullptr, 0};
        case MB: return {A[MA], 8 * b, 0, nullptr, 0};
        case D3A: return {A[MB], 8 * b, 1, A[E3B], 4 * b};
        case D3B: return {A[D3A], 4 * b, 0, nullptr, 0};
        case D2A: return {A[D3B], 4 * b, 1, A[E2B], 2 * b};
        case D2B: return {A[D2A], 2 * b, 0, nullptr, 0};
        case D1A: return {A[D2B], 2 * b, 1, A[E1B], b};
        default: return {A[D1A], b, 0, nullptr, 0};
    }
}

int forward(cageo_trainer* t) {
    const int b = t->base, n = t->n;
    void* s = t->stream;
    for (int i = 0; i < TR_NLAYERS; ++i) {
        const TrLayer& L = TR_LAYERS[i];
        const LayerIn in = layer_in(t, i);
        Param* w = find(t, std::string(L.name) + ".w");
        Param* bias = find(t, std::string(L.name) + ".b");
        CAGEO_TRY(cageo_nn_conv3x3(s, in.in1, in.c1, in.up1, in.in2, in.c2, w->pb, bias->p,
                                   t->act[i], nullptr, n, t->h >> L.level, t->w >> L.level,
                                   L.com * b, 1));
        // the encoder outputs of levels 0-2 are pooled for the next level
        if (i == E1B || i == E2B || i == E3B) {
            const int lv = L.level;
            CAGEO_TRY(cageo_nn_avgpool2(s, t->act[i], t->act[TR_NLAYERS + lv], n, t->h >> lv,
                                        t->w >> lv, L.com * b));
        }
    }
    return cageo_nn_head(s, t->act[D1B], find(t, "head.w")->p, find(t, "head.b")->p, t->logit,
                         (long long)n * t->h * t->w, b, 0);
}

// Gradients of layer i from the fp32 gradient of its output: bias and
// weight gradients (added), and with need_dx the gradient of its input
// (fp32, its concatenated channels) into dx.
int layer_back(cageo_trainer* t, int i, const float* gout, float* dx) {
    const TrLayer& L = TR_LAYERS[i];
    const int b = t->base, n = t->n, h = t->h >> L.level, w = t->w >> L.level;
    const int co = L.com * b;
    const long long P = (long long)n * h * w;
    void* s = t->stream;
    const LayerIn in = layer_in(t, i);
    Param* pw = find(t, std::string(L.name) + ".w");
    Param* pb = find(t, std::string(L.name) + ".b");
    CAGEO_TRY(cageo_nn_relu_back_bias(s, gout, t->act[i], t->gtmp, P, co, pb->g));
    CAGEO_TRY(cageo_nn_conv3x3_wgrad(s, in.in1, in.c1, in.up1, in.in2, in.c2, t->gtmp, pw->g, n,
                                     h, w, co));
    if (!dx) return CAGEO_OK;
    const int cin = in.c1 + in.c2;
    CAGEO_TRY(cageo_nn_flip_transpose(s, pw->pb, t->wt, co, cin));
    return cageo_nn_conv3x3(s, t->gtmp, co, 0, nullptr, 0, t->wt, nullptr, nullptr, dx, n, h, w,
                            cin, 0);
}

int zero(cageo_trainer* t, float* p, long long n) {
    hipLaunchKernelGGL(k_tr_zero, dim3(tr_grid(n)), dim3(256), 0, t->stream, p, n);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

// Backward pass after forward(): parameter gradients (zeroed first) and the
// batch's mean BCE added to *d_loss.
int backward(cageo_trainer* t, float* d_loss) {
    const int b = t->base, n = t->n, h = t->h, w = t->w;
    void* s = t->stream;
    auto px = [&](int lv) { return (long long)n * (h >> lv) * (w >> lv); };
    float* const* G = t->gr.data();        // G[i]: gradient of layer i's output
    float* const* DC = t->gr.data() + TR_NLAYERS;  // DC[i]: of its concatenated input (up1 layers)
    float* const* DP = t->gr.data() + 2 * TR_NLAYERS;  // DP[lv-1]: of pooled tensor lv
    for (Param& p : t->params) CAGEO_TRY(zero(t, p.g, p.n));
    Param* hw = find(t, "head.w");
    Param* hb = find(t, "head.b");
    CAGEO_TRY(cageo_nn_head_back(s, t->act[D1B], hw->p, t->logit, t->yb, t->q, G[D1B], hw->g,
                                 hb->g, d_loss, px(0), b));
    // decoder level 1 (full resolution)
    CAGEO_TRY(layer_back(t, D1B, G[D1B], G[D1A]));
    CAGEO_TRY(layer_back(t, D1A, G[D1A], DC[D1A]));                       // [2b up | b]
    CAGEO_TRY(cageo_nn_slice_add(s, DC[D1A], G[E1B], px(0), b, 3 * b, 2 * b, 0));
    CAGEO_TRY(zero(t, G[D2B], px(1) * 2 * b));
    CAGEO_TRY(cageo_nn_upsample2_back(s, DC[D1A], G[D2B], n, h, w, 2 * b, 3 * b, 0));
    // level 2
    CAGEO_TRY(layer_back(t, D2B, G[D2B], G[D2A]));
    CAGEO_TRY(layer_back(t, D2A, G[D2A], DC[D2A]));                       // [4b up | 2b]
    CAGEO_TRY(cageo_nn_slice_add(s, DC[D2A], G[E2B], px(1), 2 * b, 6 * b, 4 * b, 0));
    CAGEO_TRY(zero(t, G[D3B], px(2) * 4 * b));
    CAGEO_TRY(cageo_nn_upsample2_back(s, DC[D2A], G[D3B], n, h >> 1, w >> 1, 4 * b, 6 * b, 0));
    // level 3
    CAGEO_TRY(layer_back(t, D3B, G[D3B], G[D3A]));
    CAGEO_TRY(layer_back(t, D3A, G[D3A], DC[D3A]));                       // [8b up | 4b]
    CAGEO_TRY(cageo_nn_slice_add(s, DC[D3A], G[E3B], px(2), 4 * b, 12 * b, 8 * b, 0));
    CAGEO_TRY(zero(t, G[MB], px(3) * 8 * b));
    CAGEO_TRY(cageo_nn_upsample2_back(s, DC[D3A], G[MB], n, h >> 2, w >> 2, 8 * b, 12 * b, 0));
    // middle
    CAGEO_TRY(layer_back(t, MB, G[MB], G[MA]));
    CAGEO_TRY(layer_back(t, MA, G[MA], DP[2]));
    CAGEO_TRY(cageo_nn_avgpool2_back(s, DP[2], G[E3B], n, h >> 2, w >> 2, 4 * b));
    // encoder
    CAGEO_TRY(layer_back(t, E3B, G[E3B], G[E3A]));
    CAGEO_TRY(layer_back(t, E3A, G[E3A], DP[1]));
    CAGEO_TRY(cageo_nn_avgpool2_back(s, DP[1], G[E2B], n, h >> 1, w >> 1, 2 * b));
    CAGEO_TRY(layer_back(t, E2B, G[E2B], G[E2A]));
    CAGEO_TRY(layer_back(t, E2A, G[E2A], DP[0]));
    CAGEO_TRY(cageo_nn_avgpool2_back(s, DP[0], G[E1B], n, h, w, b));
    CAGEO_TRY(layer_back(t, E1B, G[E1B], G[E1A]));
    CAGEO_TRY(layer_back(t, E1A, G[E1A], nullptr));
    Param* w1 = find(t, "enc1a.w");
    hipLaunchKernelGGL(k_tr_zero_pad, dim3(tr_grid(9LL * b * (TR_IN_PAD - TR_IN_CH))), dim3(256),
                       0, t->stream, w1->g, 9LL * b, TR_IN_PAD, TR_IN_CH);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int adam(cageo_trainer* t, float lr) {
    ++t->step;
    for (Param& p : t->params)
        CAGEO_TRY(cageo_nn_adam(t->stream, p.p, p.g, p.m, p.v, p.pb, p.n, lr, 0.9f, 0.999f, 1e-8f,
                                t->step));
    return CAGEO_OK;
}

int gather(cageo_trainer* t, const float* d_x, const float* d_y, const int* d_sel) {
    const long long P = (long long)t->n * t->h * t->w;
    hipLaunchKernelGGL(k_tr_gather, dim3(tr_grid(P)), dim3(256), 0, t->stream, d_x, d_y, d_sel,
                       t->n, t->h, t->w, t->xb, t->yb);
    CAGEO_HIP_TRY(hipGetLastError());
    return CAGEO_OK;
}

int ensure_loop(cageo_trainer* t, int steps, long long sel_ints) {
    if (steps > t->loss_cap) {
        if (t->loss) (void)hipFree(t->loss);
        t->loss = nullptr;
        t->loss_cap = 0;
        CAGEO_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&t->loss), (size_t)steps * sizeof(float)));
        t->loss_cap = steps;
    }
    if (sel_ints > t->sel_cap) {
        if (t->sel) (void)hipFree(t->sel);
        t->sel = nullptr;
        t->sel_cap = 0;
        CAGEO_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&t->sel), (size_t)sel_ints * sizeof(int)));
        t->sel_cap = sel_ints;
    }
    return CAGEO_OK;
}

// Parameters of a base-b network: convolution weights [9][cout][cin] and
// biases per layer, then head.w [b] and head.b [1].
int make_params(cageo_trainer* t) {
    const int b = t->base;
    for (const TrLayer& L : TR_LAYERS) {
        const int cin = L.c1m == 0 ? TR_IN_PAD : (L.c1m + L.c2m) * b, co = L.com * b;
        Param w, bias;
        w.name = std::string(L.name) + ".w";
        w.shape = {9, co, cin};
        bias.name = std::string(L.name) + ".b";
        bias.shape = {co};
        t->params.push_back(w);
        t->params.push_back(bias);
    }
    Param hw, hb;
    hw.name = "head.w";
    hw.shape = {b};
    hb.name = "head.b";
    hb.shape = {1};
    t->params.push_back(hw);
    t->params.push_back(hb);
    for (Param& p : t->params) {
        p.n = 1;
        for (int d : p.shape) p.n *= d;
        const size_t bytes = (size_t)p.n * sizeof(float);
        if (hipMalloc(reinterpret_cast<void**>(&p.p), bytes) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&p.g), bytes) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&p.m), bytes) != hipSuccess ||
            hipMalloc(reinterpret_cast<void**>(&p.v), bytes) != hipSuccess)
            return CAGEO_ERR_NOMEM;
        CAGEO_HIP_TRY(hipMemset(p.m, 0, bytes));
        CAGEO_HIP_TRY(hipMemset(p.v, 0, bytes));
        CAGEO_HIP_TRY(hipMemset(p.g, 0, bytes));
        if (p.shape.size() == 3 &&
            hipMalloc(reinterpret_cast<void**>(&p.pb), (size_t)p.n * sizeof(uint16_t)) != hipSuccess)
            return CAGEO_ERR_NOMEM;
    }
    return CAGEO_OK;
}

// Host values of every parameter -> device (bf16 copies refreshed).
int upload(cageo_trainer* t, const std::vector<std::vector<float>>& host) {
    for (size_t i = 0; i < t->params.size(); ++i) {
        Param& p = t->params[i];
        CAGEO_HIP_TRY(hipMemcpy(p.p, host[i].data(), (size_t)p.n * sizeof(float),
                                hipMemcpyHostToDevice));
        if (p.pb) CAGEO_TRY(cageo_nn_to_bf16(t->stream, p.p, p.pb, p.n));
    }
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    return CAGEO_OK;
}

// He-normal weights (fan-in 9 x used input channels; the padded input
// channels of the first layer zero), zero biases, head N(0, 1/b).
void init_host(cageo_trainer* t, uint64_t seed, std::vector<std::vector<float>>& host) {
    std::mt19937_64 rng(seed);
    for (const Param& p : t->params) {
        std::vector<float> v((size_t)p.n, 0.0f);
        if (p.shape.size() == 3) {
            const int co = p.shape[1], cin = p.shape[2];
            const bool first = p.name == "enc1a.w";
            std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / (9.0 * (first ? TR_IN_CH : cin))));
            for (int tap = 0; tap < 9; ++tap)
                for (int o = 0; o < co; ++o)
                    for (int c = 0; c < cin; ++c)
                        v[((size_t)tap * co + o) * cin + c] = (!first || c < TR_IN_CH) ? (float)nd(rng) : 0.0f;
        } else if (p.name == "head.w") {
            std::normal_distribution<double> nd(0.0, std::sqrt(1.0 / t->base));
            for (float& x : v) x = (float)nd(rng);
        }
        host.push_back(std::move(v));
    }
}

// Parameters from a weights file of the surrogate's format (convolution
// weights bf16, the rest fp32); the network's base must match.
int read_weights(cageo_trainer* t, const char* path, std::vector<std::vector<float>>& host) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return CAGEO_ERR_ARG;
    int st = CAGEO_OK;
    char magic[8];
    int32_t hdr[2];
    host.assign(t->params.size(), {});
    if (!tr_read(f, magic, 8) || std::memcmp(magic, "CGSURR01", 8) != 0 ||
        !tr_read(f, hdr, sizeof(hdr)) || hdr[0] != t->base || hdr[1] <= 0 || hdr[1] > 256)
        st = CAGEO_ERR_ARG;
    for (int k = 0; st == CAGEO_OK && k < hdr[1]; ++k) {
        int32_t ln = 0, meta[2];
        if (!tr_read(f, &ln, 4) || ln <= 0 || ln > 64) { st = CAGEO_ERR_ARG; break; }
        std::string name(ln, '\0');
        if (!tr_read(f, &name[0], ln) || !tr_read(f, meta, 8) || meta[1] < 1 || meta[1] > 4 ||
            (meta[0] != 0 && meta[0] != 1)) { st = CAGEO_ERR_ARG; break; }
        std::vector<int> dims(meta[1]);
        long long nel = 1;
        for (int d = 0; d < meta[1]; ++d) {
            int32_t v = 0;
            if (!tr_read(f, &v, 4) || v <= 0 || v > (1 << 20)) { st = CAGEO_ERR_ARG; break; }
            dims[d] = v;
            nel *= v;
        }
        if (st != CAGEO_OK) break;
        std::vector<float> vals((size_t)nel);
        if (meta[0] == 1) {
            std::vector<uint16_t> raw((size_t)nel);
            if (!tr_read(f, raw.data(), (size_t)nel * 2)) { st = CAGEO_ERR_ARG; break; }
            for (long long i = 0; i < nel; ++i) {
                const uint32_t u = (uint32_t)raw[(size_t)i] << 16;
                std::memcpy(&vals[(size_t)i], &u, 4);
            }
        } else if (!tr_read(f, vals.data(), (size_t)nel * 4)) {
            st = CAGEO_ERR_ARG;
            break;
        }
        for (size_t i = 0; i < t->params.size(); ++i)
            if (t->params[i].name == name) {
                if (t->params[i].shape != dims) st = CAGEO_ERR_ARG;
                else host[i] = std::move(vals);
                break;
            }
    }
    std::fclose(f);
    for (size_t i = 0; st == CAGEO_OK && i < host.size(); ++i)
        if ((long long)host[i].size() != t->params[i].n) st = CAGEO_ERR_ARG;  // a tensor is missing
    return st;
}

}  // namespace

}  // namespace cageo

using cageo::Param;

extern "C" {

int cageo_trainer_destroy(cageo_trainer* t) {
    if (!t) return CAGEO_ERR_ARG;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    else (void)hipDeviceSynchronize();
    cageo::free_batch(t);
    for (Param& p : t->params) {
        for (float* q : {p.p, p.g, p.m, p.v})
            if (q) (void)hipFree(q);
        if (p.pb) (void)hipFree(p.pb);
    }
    if (t->loss) (void)hipFree(t->loss);
    if (t->sel) (void)hipFree(t->sel);
    delete t;
    return CAGEO_OK;
}

int cageo_trainer_create(cageo_trainer** out, int base, uint64_t seed, const char* init_weights,
                         void* hip_stream) {
    if (!out || base <= 0 || base % 32 || base > 128) return CAGEO_ERR_ARG;
    cageo_trainer* t = new cageo_trainer();
    t->stream = reinterpret_cast<hipStream_t>(hip_stream);
    t->base = base;
    int st = cageo::make_params(t);
    std::vector<std::vector<float>> host;
    if (st == CAGEO_OK) {
        if (init_weights) st = cageo::read_weights(t, init_weights, host);
        else cageo::init_host(t, seed, host);
    }
    if (st == CAGEO_OK) st = cageo::upload(t, host);
    if (st != CAGEO_OK) {
        cageo_trainer_destroy(t);
        return st;
    }
    *out = t;
    return CAGEO_OK;
}

int cageo_trainer_gradients(cageo_trainer* t, const float* d_x, const float* d_target, int n,
                            int h, int w, float* loss) {
    if (!t || !d_x || !d_target || n <= 0 || h <= 0 || w <= 0 || h % 8 || w % 8)
        return CAGEO_ERR_ARG;
    CAGEO_TRY(cageo::ensure_batch(t, n, h, w));
    CAGEO_TRY(cageo::ensure_loop(t, 1, 2LL * n));
    std::vector<int> sel(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        sel[2 * (size_t)i] = i;
        sel[2 * (size_t)i + 1] = 0;
    }
    CAGEO_HIP_TRY(hipMemcpyAsync(t->sel, sel.data(), sel.size() * sizeof(int),
                                 hipMemcpyHostToDevice, t->stream));
    CAGEO_TRY(cageo::zero(t, t->loss, 1));
    CAGEO_TRY(cageo::gather(t, d_x, d_target, t->sel));
    CAGEO_TRY(cageo::forward(t));
    CAGEO_TRY(cageo::backward(t, t->loss));
    float l = 0.0f;
    CAGEO_HIP_TRY(hipMemcpyAsync(&l, t->loss, sizeof(float), hipMemcpyDeviceToHost, t->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    if (loss) *loss = l;
    return CAGEO_OK;
}

int cageo_trainer_step(cageo_trainer* t, const float* d_x, const float* d_target, int n, int h,
                       int w, float lr, float* loss) {
    CAGEO_TRY(cageo_trainer_gradients(t, d_x, d_target, n, h, w, loss));
    CAGEO_TRY(cageo::adam(t, lr));
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    return CAGEO_OK;
}

int cageo_surrogate_train(cageo_trainer* t, const float* d_x, const float* d_y, int ntiles, int h,
                          int w, int batch, float lr, float lr_floor, int first_step, int nsteps,
                          int total_steps, uint64_t seed, float* loss_hist) {
    if (!t || !d_x || !d_y || ntiles <= 0 || h <= 0 || w <= 0 || h % 8 || w % 8 || batch <= 0 ||
        nsteps < 0 || first_step < 0 || total_steps < first_step + nsteps || !(lr > 0.0f))
        return CAGEO_ERR_ARG;
    if (nsteps == 0) return CAGEO_OK;
    CAGEO_TRY(cageo::ensure_batch(t, batch, h, w));
    CAGEO_TRY(cageo::ensure_loop(t, nsteps, 2LL * batch * nsteps));
    // every step's batch, drawn from the step number (a run in pieces draws
    // what one run would)
    std::vector<int> sel(2 * (size_t)batch * nsteps);
    for (int k = 0; k < nsteps; ++k) {
        uint64_t s = seed ^ (0x5851f42d4c957f2dull * (uint64_t)(first_step + k + 1));
        for (int i = 0; i < batch; ++i) {
            const uint64_t r = cageo::splitmix64(s);
            const size_t o = 2 * ((size_t)k * batch + i);
            sel[o] = (int)(r % (uint64_t)ntiles);
            const int tf = (int)((r >> 40) & 7u);
            sel[o + 1] = h == w ? tf : (tf & 6);  // no quarter turns of a non-square tile
        }
    }
    CAGEO_HIP_TRY(hipMemcpyAsync(t->sel, sel.data(), sel.size() * sizeof(int),
                                 hipMemcpyHostToDevice, t->stream));
    CAGEO_TRY(cageo::zero(t, t->loss, nsteps));
    for (int k = 0; k < nsteps; ++k) {
        // linear warm-up over the first 2 % of the run (at most 200 steps),
        // then cosine decay from lr, not below lr * lr_floor
        const int step = first_step + k;
        const int warm = std::max(1, std::min(200, total_steps / 50));
        const double ph = (double)step / (double)total_steps;
        float cur = std::max((float)(lr * 0.5 * (1.0 + std::cos(M_PI * ph))), lr * lr_floor);
        if (step < warm) cur = lr * (float)(step + 1) / (float)warm;
        CAGEO_TRY(cageo::gather(t, d_x, d_y, t->sel + 2 * (size_t)k * batch));
        CAGEO_TRY(cageo::forward(t));
        CAGEO_TRY(cageo::backward(t, t->loss + k));
        CAGEO_TRY(cageo::adam(t, cur));
    }
    if (loss_hist)
        CAGEO_HIP_TRY(hipMemcpyAsync(loss_hist, t->loss, (size_t)nsteps * sizeof(float),
                                     hipMemcpyDeviceToHost, t->stream));
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    return CAGEO_OK;
}

int cageo_trainer_read(cageo_trainer* t, const char* name, int what, float* host, long long n) {
    if (!t || !name || !host) return CAGEO_ERR_ARG;
    Param* p = cageo::find(t, name);
    if (!p || n != p->n || what < 0 || what > 1) return CAGEO_ERR_ARG;
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    CAGEO_HIP_TRY(hipMemcpy(host, what == 0 ? p->p : p->g, (size_t)n * sizeof(float),
                            hipMemcpyDeviceToHost));
    return CAGEO_OK;
}

int cageo_trainer_save(cageo_trainer* t, const char* path) {
    if (!t || !path) return CAGEO_ERR_ARG;
    CAGEO_HIP_TRY(hipStreamSynchronize(t->stream));
    FILE* f = std::fopen(path, "wb");
    if (!f) return CAGEO_ERR_ARG;
    bool ok = std::fwrite("CGSURR01", 1, 8, f) == 8;
    const int32_t hdr[2] = {t->base, (int32_t)t->params.size()};
    ok = ok && std::fwrite(hdr, 4, 2, f) == 2;
    for (const Param& p : t->params) {
        if (!ok) break;
        std::vector<float> v((size_t)p.n);
        if (hipMemcpy(v.data(), p.p, (size_t)p.n * sizeof(float), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            ok = false;
            break;
        }
        const bool bf = p.shape.size() == 3;
        const int32_t ln = (int32_t)p.name.size(), meta[2] = {bf ? 1 : 0, (int32_t)p.shape.size()};
        ok = std::fwrite(&ln, 4, 1, f) == 1 && std::fwrite(p.name.data(), 1, ln, f) == (size_t)ln &&
             std::fwrite(meta, 4, 2, f) == 2 &&
             std::fwrite(p.shape.data(), 4, p.shape.size(), f) == p.shape.size();
        if (!ok) break;
        if (bf) {
            std::vector<uint16_t> raw((size_t)p.n);
            for (size_t i = 0; i < raw.size(); ++i) {  // round to nearest even
                uint32_t u;
                std::memcpy(&u, &v[i], 4);
                raw[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
            }
            ok = std::fwrite(raw.data(), 2, raw.size(), f) == raw.size();
        } else {
            ok = std::fwrite(v.data(), 4, v.size(), f) == v.size();
        }
    }
    ok = std::fclose(f) == 0 && ok;
    return ok ? CAGEO_OK : CAGEO_ERR_ARG;
}

}  // extern "C"
