// surrogate.hip -- C ABI of include/lbm_surrogate.h: the 3D U-Net surrogate
// (forward, backward, Adam training, prediction) on the BF16 MFMA kernels of
// surrogate_kernels.hpp, and the warm start of a solver from its prediction.
//
// Network (S^3 input, S a multiple of 8):
//   level 0 (S):   conv 8->16, conv 16->16 = s1          pool
//   level 1 (S/2): conv 16->32, conv 32->32 = s2         pool
//   level 2 (S/4): conv 32->64, conv 64->64 = s3         pool
//   level 3 (S/8): conv 64->128, conv 128->128
//   up to S/4, concat s3: conv 192->64, conv 64->64
//   up to S/2, concat s2: conv 96->32,  conv 32->32
//   up to S,   concat s1: conv 48->16,  conv 16->16
//   head: 1x1x1 conv 16->8 (outputs 0-3 = u/U, v/U, w/U, p/U^2; 4-7 unused)
// 3x3x3 convolutions with zero padding and ReLU, 2x2x2 average pooling,
// nearest-neighbour upsampling.  Input channels: 3 used (lid mask, c_Re, c_N),
// padded to 8 with zeros (the MFMA operand loads take 8 channels at a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "lbm_surrogate.h"
#include "lbm_internal.hpp"
#include "surrogate_kernels.hpp"

using namespace sur;
using lbm::ArgError;
using lbm::HipError;

namespace {

struct ConvSpec {
    int cin, cout, T, relu, level;
};
constexpr int kNConv = 15;
const ConvSpec kConv[kNConv] = {
    {8, 16, 27, 1, 0},    {16, 16, 27, 1, 0},   {16, 32, 27, 1, 1},  {32, 32, 27, 1, 1},  {32, 64, 27, 1, 2},
    {64, 64, 27, 1, 2},   {64, 128, 27, 1, 3},  {128, 128, 27, 1, 3}, {192, 64, 27, 1, 2}, {64, 64, 27, 1, 2},
    {96, 32, 27, 1, 1},   {32, 32, 27, 1, 1},   {48, 16, 27, 1, 0},  {16, 16, 27, 1, 0},  {16, 8, 1, 0, 0}};
constexpr int kInUsed = 3, kInPad = 8, kOutUsed = 4;

// Activation tensors (BF16, channels last) and their levels / channels.
enum Tensor { X0, A1, S1, P1, A2, S2, P2, A3, S3, P3, B1, B2, U3, D3A, D3B, U2, D2A, D2B, U1, D1A, D1B, NT };
const int kTLevel[NT] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0};
const int kTChan[NT] = {8, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128, 128, 192, 64, 64, 96, 32, 32, 48, 16, 16};
// conv i reads kConvIn[i] and writes kConvOut[i] (the head writes the FP32 output)
const int kConvIn[kNConv] = {X0, A1, P1, A2, P2, A3, P3, B1, U3, D3A, U2, D2A, U1, D1A, D1B};
const int kConvOut[kNConv] = {A1, S1, A2, S2, A3, S3, B1, B2, D3A, D3B, D2A, D2B, D1A, D1B, -1};

constexpr int kLossBlocks = 256;
const float kBeta1 = 0.9f, kBeta2 = 0.999f, kEps = 1e-8f;

thread_local std::string g_sur_create_error;

inline int grid_for(int64_t n) { return (int)std::min<int64_t>(4096, (n + 255) / 256); }

}  // namespace

struct lbm_sur {
    int device = 0, S = 32;
    hipStream_t stream = nullptr;
    int64_t woff[kNConv] = {}, boff[kNConv] = {}, nparam = 0;
    float *w = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wb[kNConv] = {}, *wt[kNConv] = {};
    int64_t t_adam = 0;
    // per-batch buffers for `cap` samples
    int cap = 0;
    bf16 *act[NT] = {};
    float *dact[NT] = {};
    bf16 *gpre[kNConv] = {};
    float *out = nullptr, *tgt = nullptr, *stage = nullptr, *part = nullptr;
    int64_t part_len = 0;
    double *lpart = nullptr;
    // resident training set
    int nd = 0;
    bf16 *din = nullptr;
    float *dtg = nullptr;
    std::string err;

    int side(int level) const { return S >> level; }
    int64_t points(int level, int B) const {
        const int64_t s = side(level);
        return (int64_t)B * s * s * s;
    }
    int64_t wlen(int i) const { return (int64_t)kConv[i].cout * kConv[i].T * kConv[i].cin; }

    void free_batch() {
        for (int i = 0; i < NT; i++) {
            if (act[i]) (void)hipFree(act[i]);
            if (dact[i]) (void)hipFree(dact[i]);
            act[i] = nullptr;
            dact[i] = nullptr;
        }
        for (int i = 0; i < kNConv; i++) {
            if (gpre[i]) (void)hipFree(gpre[i]);
            gpre[i] = nullptr;
        }
        for (void *p : {(void *)out, (void *)tgt, (void *)stage, (void *)part, (void *)lpart})
            if (p) (void)hipFree(p);
        out = tgt = stage = part = nullptr;
        lpart = nullptr;
        cap = 0;
    }
    void release() {
        free_batch();
        for (int i = 0; i < kNConv; i++) {
            if (wb[i]) (void)hipFree(wb[i]);
            if (wt[i]) (void)hipFree(wt[i]);
            wb[i] = wt[i] = nullptr;
        }
        for (void *p : {(void *)w, (void *)grad, (void *)m, (void *)v, (void *)din, (void *)dtg})
            if (p) (void)hipFree(p);
        w = grad = m = v = nullptr;
        din = nullptr;
        dtg = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

// ---------------------------------------------------------------- launches
int nb_for(int n) { return n >= 64 ? 4 : (n >= 32 ? 2 : 1); }

template <int MODE>
void launch_points(const ConvArgs &a, int N, hipStream_t s) {
    const int nb = nb_for(N);
    const dim3 grid((unsigned)((a.P + 63) / 64), (unsigned)((N + 16 * nb - 1) / (16 * nb)));
    if (nb == 4)
        hipLaunchKernelGGL((conv_points_kernel<MODE, 4>), grid, dim3(256), 0, s, a);
    else if (nb == 2)
        hipLaunchKernelGGL((conv_points_kernel<MODE, 2>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((conv_points_kernel<MODE, 1>), grid, dim3(256), 0, s, a);
    HIPX(hipGetLastError());
}

ConvArgs conv_args(const lbm_sur *s, int i, int B) {
    ConvArgs a;
    std::memset(&a, 0, sizeof a);
    a.S = s->side(kConv[i].level);
    a.P = (int)s->points(kConv[i].level, B);
    a.Cin = kConv[i].cin;
    a.Cout = kConv[i].cout;
    a.T = kConv[i].T;
    a.relu = kConv[i].relu;
    return a;
}

void forward(lbm_sur *s, int B) {
    hipStream_t st = s->stream;
    for (int i = 0; i < kNConv; i++) {
        // inputs that are pooled or concatenated first
        const int in = kConvIn[i];
        if (in == P1 || in == P2 || in == P3) {
            const int src = in == P1 ? S1 : in == P2 ? S2 : S3;
            const int lv = kTLevel[src];
            const int64_t n = s->points(lv + 1, B) * kTChan[in];
            hipLaunchKernelGGL(pool_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, st, s->act[src], s->side(lv), B,
                               kTChan[src], s->act[in]);
            HIPX(hipGetLastError());
        } else if (in == U3 || in == U2 || in == U1) {
            const int low = in == U3 ? B2 : in == U2 ? D3B : D2B;
            const int skip = in == U3 ? S3 : in == U2 ? S2 : S1;
            const int64_t n = s->points(kTLevel[in], B) * kTChan[in];
            hipLaunchKernelGGL(upcat_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, st, s->act[low], s->act[skip],
                               s->side(kTLevel[in]), B, kTChan[low], kTChan[skip], s->act[in]);
            HIPX(hipGetLastError());
        }
        ConvArgs a = conv_args(s, i, B);
        a.x = s->act[in];
        a.w = s->wb[i];
        a.bias = s->w + s->boff[i];
        if (kConvOut[i] >= 0)
            a.y = s->act[kConvOut[i]];
        else
            a.yf = s->out;
        launch_points<FWD>(a, a.Cout, st);
    }
}

// Loss of the current forward pass against s->tgt; writes the head's G when g.
double loss_and_head_grad(lbm_sur *s, int B, bool g) {
    const int64_t P = s->points(0, B);
    hipLaunchKernelGGL(mse_kernel, dim3(kLossBlocks), dim3(256), 0, s->stream, s->out, s->tgt, P,
                       g ? s->gpre[kNConv - 1] : nullptr, s->lpart);
    HIPX(hipGetLastError());
    double h[kLossBlocks];
    HIPX(hipMemcpyAsync(h, s->lpart, sizeof h, hipMemcpyDeviceToHost, s->stream));
    HIPX(hipStreamSynchronize(s->stream));
    double sum = 0.0;
    for (double v : h) sum += v;
    return sum / (double)(kOutUsed * P);
}

void wgrad(lbm_sur *s, int i, int B) {
    ConvArgs a = conv_args(s, i, B);
    a.x = s->act[kConvIn[i]];
    a.g = s->gpre[i];
    const int N = a.T * a.Cin;
    const int nb = N >= 64 ? 4 : 1;
    const int waves = ((a.Cout + 15) / 16) * ((N + 16 * nb - 1) / (16 * nb));
    const int target = std::max(1, 2048 / waves);
    int kchunk = (int)((a.P + target - 1) / target);
    kchunk = std::max(256, (kchunk + 31) / 32 * 32);
    const int nchunk = (a.P + kchunk - 1) / kchunk;
    const int64_t len = (int64_t)a.Cout * N;
    if ((int64_t)nchunk * len > s->part_len) throw ArgError{LBM_ERR_STATE, "weight-gradient scratch too small"};
    a.kchunk = kchunk;
    a.part = s->part;
    const dim3 grid((unsigned)waves, (unsigned)nchunk);
    if (nb == 4)
        hipLaunchKernelGGL((conv_wgrad_kernel<4>), grid, dim3(64), 0, s->stream, a);
    else
        hipLaunchKernelGGL((conv_wgrad_kernel<1>), grid, dim3(64), 0, s->stream, a);
    HIPX(hipGetLastError());
    hipLaunchKernelGGL(sum_chunks_kernel, dim3(grid_for(len)), dim3(256), 0, s->stream, s->part, nchunk, len,
                       s->grad + s->woff[i]);
    HIPX(hipGetLastError());
    hipLaunchKernelGGL(bias_grad_kernel, dim3(a.Cout), dim3(256), 0, s->stream, s->gpre[i], a.P, a.Cout,
                       s->grad + s->boff[i]);
    HIPX(hipGetLastError());
}

void dgrad(lbm_sur *s, int i, int B, float *dx) {
    ConvArgs a = conv_args(s, i, B);
    a.g = s->gpre[i];
    a.w = s->wt[i];
    a.yf = dx;
    launch_points<DGRAD>(a, a.Cin, s->stream);
}

void relu_grad(lbm_sur *s, int i, int B) {
    const int t = kConvOut[i];
    const int64_t n = s->points(kTLevel[t], B) * kTChan[t];
    hipLaunchKernelGGL(relu_grad_kernel, dim3(grid_for(n)), dim3(256), 0, s->stream, s->dact[t], s->act[t], n,
                       s->gpre[i]);
    HIPX(hipGetLastError());
}

void pool_bwd(lbm_sur *s, int pooled, int full, int B) {
    const int lv = kTLevel[full];
    const int64_t n = s->points(lv, B) * kTChan[full];
    hipLaunchKernelGGL(pool_bwd_add_kernel, dim3(grid_for(n)), dim3(256), 0, s->stream, s->dact[pooled],
                       s->side(lv), B, kTChan[full], s->dact[full]);
    HIPX(hipGetLastError());
}

void upcat_bwd(lbm_sur *s, int cat, int low, int skip, int B) {
    const int lv = kTLevel[cat];
    const int64_t n = s->points(lv + 1, B) * kTChan[low] + s->points(lv, B) * kTChan[skip];
    hipLaunchKernelGGL(upcat_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, s->stream, s->dact[cat], s->side(lv), B,
                       kTChan[low], kTChan[skip], s->dact[low], s->dact[skip]);
    HIPX(hipGetLastError());
}

// Gradient of the loss (head G already formed) into s->grad.
void backward(lbm_sur *s, int B) {
    // head and decoder level 0
    wgrad(s, 14, B);
    dgrad(s, 14, B, s->dact[D1B]);
    relu_grad(s, 13, B);
    wgrad(s, 13, B);
    dgrad(s, 13, B, s->dact[D1A]);
    relu_grad(s, 12, B);
    wgrad(s, 12, B);
    dgrad(s, 12, B, s->dact[U1]);
    upcat_bwd(s, U1, D2B, S1, B);
    // decoder level 1
    relu_grad(s, 11, B);
    wgrad(s, 11, B);
    dgrad(s, 11, B, s->dact[D2A]);
    relu_grad(s, 10, B);
    wgrad(s, 10, B);
    dgrad(s, 10, B, s->dact[U2]);
    upcat_bwd(s, U2, D3B, S2, B);
    // decoder level 2
    relu_grad(s, 9, B);
    wgrad(s, 9, B);
    dgrad(s, 9, B, s->dact[D3A]);
    relu_grad(s, 8, B);
    wgrad(s, 8, B);
    dgrad(s, 8, B, s->dact[U3]);
    upcat_bwd(s, U3, B2, S3, B);
    // bottleneck
    relu_grad(s, 7, B);
    wgrad(s, 7, B);
    dgrad(s, 7, B, s->dact[B1]);
    relu_grad(s, 6, B);
    wgrad(s, 6, B);
    dgrad(s, 6, B, s->dact[P3]);
    pool_bwd(s, P3, S3, B);
    // encoder level 2
    relu_grad(s, 5, B);
    wgrad(s, 5, B);
    dgrad(s, 5, B, s->dact[A3]);
    relu_grad(s, 4, B);
    wgrad(s, 4, B);
    dgrad(s, 4, B, s->dact[P2]);
    pool_bwd(s, P2, S2, B);
    // encoder level 1
    relu_grad(s, 3, B);
    wgrad(s, 3, B);
    dgrad(s, 3, B, s->dact[A2]);
    relu_grad(s, 2, B);
    wgrad(s, 2, B);
    dgrad(s, 2, B, s->dact[P1]);
    pool_bwd(s, P1, S1, B);
    // encoder level 0
    relu_grad(s, 1, B);
    wgrad(s, 1, B);
    dgrad(s, 1, B, s->dact[A1]);
    relu_grad(s, 0, B);
    wgrad(s, 0, B);
}

void cast_all(lbm_sur *s) {
    for (int i = 0; i < kNConv; i++) {
        const int64_t n = s->wlen(i);
        hipLaunchKernelGGL(cast_weights_kernel, dim3(grid_for(n)), dim3(256), 0, s->stream, s->w + s->woff[i],
                           kConv[i].cout, kConv[i].T, kConv[i].cin, s->wb[i], s->wt[i]);
        HIPX(hipGetLastError());
    }
}

void ensure_batch(lbm_sur *s, int B) {
    if (B <= s->cap) return;
    s->free_batch();
    for (int t = 0; t < NT; t++) {
        const int64_t n = s->points(kTLevel[t], B) * kTChan[t];
        HIPX(hipMalloc(&s->act[t], sizeof(bf16) * n));
        if (t != X0) HIPX(hipMalloc(&s->dact[t], sizeof(float) * n));
    }
    int64_t plen = 0;
    for (int i = 0; i < kNConv; i++) {
        const int64_t P = s->points(kConv[i].level, B);
        HIPX(hipMalloc(&s->gpre[i], sizeof(bf16) * P * kConv[i].cout));
        // largest split-K scratch any weight gradient asks for (see wgrad)
        const int N = kConv[i].T * kConv[i].cin, nb = N >= 64 ? 4 : 1;
        const int waves = ((kConv[i].cout + 15) / 16) * ((N + 16 * nb - 1) / (16 * nb));
        const int target = std::max(1, 2048 / waves);
        int kchunk = (int)((P + target - 1) / target);
        kchunk = std::max(256, (kchunk + 31) / 32 * 32);
        const int64_t nchunk = (P + kchunk - 1) / kchunk;
        plen = std::max(plen, nchunk * kConv[i].cout * N);
    }
    const int64_t P0 = s->points(0, B);
    HIPX(hipMalloc(&s->out, sizeof(float) * P0 * 8));
    HIPX(hipMalloc(&s->tgt, sizeof(float) * P0 * kOutUsed));
    HIPX(hipMalloc(&s->stage, sizeof(float) * P0 * kInUsed));
    HIPX(hipMalloc(&s->part, sizeof(float) * plen));
    HIPX(hipMalloc(&s->lpart, sizeof(double) * kLossBlocks));
    s->part_len = plen;
    s->cap = B;
}

// Host inputs [B][S^3][3] -> packed BF16 input tensor.
void upload_inputs(lbm_sur *s, int B, const float *inputs) {
    const int64_t P = s->points(0, B);
    HIPX(hipMemcpyAsync(s->stage, inputs, sizeof(float) * P * kInUsed, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for(P * kInPad)), dim3(256), 0, s->stream, s->stage, P, kInUsed,
                       kInPad, s->act[X0]);
    HIPX(hipGetLastError());
}

template <typename F>
int guarded(lbm_sur *s, F &&f) {
    if (!s) return LBM_ERR_ARG;
    try {
        HIPX(hipSetDevice(s->device));
        f();
        return LBM_OK;
    } catch (const HipError &e) {
        s->err = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
        return LBM_ERR_HIP;
    } catch (const ArgError &e) {
        s->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        s->err = "host allocation failed";
        return LBM_ERR_NOMEM;
    }
}

}  // namespace

extern "C" {

const char *lbm_sur_create_error(void) { return g_sur_create_error.c_str(); }

int lbm_sur_create(int device, int S, uint64_t seed, lbm_sur **out) {
    g_sur_create_error.clear();
    if (!out || S < 8 || S % 8) {
        g_sur_create_error = "S must be a positive multiple of 8";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    lbm_sur *s = new (std::nothrow) lbm_sur();
    if (!s) return LBM_ERR_NOMEM;
    s->device = device;
    s->S = S;
    int64_t off = 0;
    for (int i = 0; i < kNConv; i++) {
        s->woff[i] = off;
        off += s->wlen(i);
    }
    for (int i = 0; i < kNConv; i++) {
        s->boff[i] = off;
        off += kConv[i].cout;
    }
    s->nparam = off;
    // He-normal weights (fan-in over the channels in use), zero biases; the
    // head starts small, its unused outputs and the padded inputs at zero.
    std::vector<float> h((size_t)off, 0.f);
    std::mt19937_64 rng(seed);
    for (int i = 0; i < kNConv; i++) {
        const ConvSpec &c = kConv[i];
        const int cin_used = i == 0 ? kInUsed : c.cin;
        const int cout_used = i == kNConv - 1 ? kOutUsed : c.cout;
        const double sd = std::sqrt(2.0 / (c.T * cin_used)) * (i == kNConv - 1 ? 0.1 : 1.0);
        std::normal_distribution<double> nd(0.0, sd);
        for (int co = 0; co < c.cout; co++)
            for (int t = 0; t < c.T; t++)
                for (int ci = 0; ci < c.cin; ci++) {
                    const double r = nd(rng);
                    if (co < cout_used && ci < cin_used) h[(size_t)(s->woff[i] + ((int64_t)co * c.T + t) * c.cin + ci)] = (float)r;
                }
    }
    try {
        HIPX(hipSetDevice(device));
        HIPX(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (float **p : {&s->w, &s->grad, &s->m, &s->v}) {
            HIPX(hipMalloc(p, sizeof(float) * off));
            HIPX(hipMemsetAsync(*p, 0, sizeof(float) * off, s->stream));
        }
        HIPX(hipMemcpyAsync(s->w, h.data(), sizeof(float) * off, hipMemcpyHostToDevice, s->stream));
        for (int i = 0; i < kNConv; i++) {
            HIPX(hipMalloc(&s->wb[i], sizeof(bf16) * s->wlen(i)));
            HIPX(hipMalloc(&s->wt[i], sizeof(bf16) * s->wlen(i)));
        }
        cast_all(s);
        HIPX(hipStreamSynchronize(s->stream));
    } catch (const HipError &e) {
        g_sur_create_error = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
        s->release();
        delete s;
        return LBM_ERR_HIP;
    }
    *out = s;
    return LBM_OK;
}

int lbm_sur_destroy(lbm_sur *s) {
    if (!s) return LBM_ERR_ARG;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->release();
    delete s;
    return LBM_OK;
}

const char *lbm_sur_last_error(const lbm_sur *s) { return s ? s->err.c_str() : "null surrogate"; }

int64_t lbm_sur_num_params(const lbm_sur *s) { return s ? s->nparam : LBM_ERR_ARG; }

int lbm_sur_get_params(lbm_sur *s, float *out) {
    return guarded(s, [&] {
        if (!out) throw ArgError{LBM_ERR_ARG, "null output"};
        HIPX(hipMemcpyAsync(out, s->w, sizeof(float) * s->nparam, hipMemcpyDeviceToHost, s->stream));
        HIPX(hipStreamSynchronize(s->stream));
    });
}

int lbm_sur_set_params(lbm_sur *s, const float *in) {
    return guarded(s, [&] {
        if (!in) throw ArgError{LBM_ERR_ARG, "null input"};
        HIPX(hipMemcpyAsync(s->w, in, sizeof(float) * s->nparam, hipMemcpyHostToDevice, s->stream));
        cast_all(s);
        HIPX(hipStreamSynchronize(s->stream));
    });
}

int lbm_sur_encode(int S, double Re, int N, float *inputs) {
    if (S < 2 || !inputs || !(Re > 0) || N < 2) return LBM_ERR_ARG;
    const float cre = (float)(std::log(Re / 800.0) / std::log(4.0));
    const float cn = (float)std::log2(N / 64.0);
    for (int z = 0; z < S; z++)
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                float *p = inputs + 3 * (((int64_t)z * S + y) * S + x);
                p[0] = y == S - 1 ? 1.f : 0.f;
                p[1] = cre;
                p[2] = cn;
            }
    return LBM_OK;
}

int lbm_sur_gradient(lbm_sur *s, int n, const float *inputs, const float *targets, double *loss, float *grad) {
    return guarded(s, [&] {
        if (n <= 0 || !inputs || !targets) throw ArgError{LBM_ERR_ARG, "bad arguments"};
        ensure_batch(s, n);
        upload_inputs(s, n, inputs);
        HIPX(hipMemcpyAsync(s->tgt, targets, sizeof(float) * s->points(0, n) * kOutUsed, hipMemcpyHostToDevice,
                            s->stream));
        forward(s, n);
        const double l = loss_and_head_grad(s, n, true);
        backward(s, n);
        if (grad) HIPX(hipMemcpyAsync(grad, s->grad, sizeof(float) * s->nparam, hipMemcpyDeviceToHost, s->stream));
        HIPX(hipStreamSynchronize(s->stream));
        if (loss) *loss = l;
    });
}

int lbm_sur_load_data(lbm_sur *s, int n, const float *inputs, const float *targets) {
    return guarded(s, [&] {
        if (n <= 0 || !inputs || !targets) throw ArgError{LBM_ERR_ARG, "bad arguments"};
        if (s->din) (void)hipFree(s->din);
        if (s->dtg) (void)hipFree(s->dtg);
        s->din = nullptr;
        s->dtg = nullptr;
        s->nd = 0;
        const int64_t ps = s->points(0, 1);
        HIPX(hipMalloc(&s->din, sizeof(bf16) * ps * kInPad * n));
        HIPX(hipMalloc(&s->dtg, sizeof(float) * ps * kOutUsed * n));
        ensure_batch(s, 1);
        for (int i = 0; i < n; i++) {
            HIPX(hipMemcpyAsync(s->stage, inputs + (int64_t)i * ps * kInUsed, sizeof(float) * ps * kInUsed,
                                hipMemcpyHostToDevice, s->stream));
            hipLaunchKernelGGL(pack_input_kernel, dim3(grid_for(ps * kInPad)), dim3(256), 0, s->stream, s->stage, ps,
                               kInUsed, kInPad, s->din + (int64_t)i * ps * kInPad);
            HIPX(hipGetLastError());
            HIPX(hipStreamSynchronize(s->stream));
        }
        HIPX(hipMemcpyAsync(s->dtg, targets, sizeof(float) * ps * kOutUsed * n, hipMemcpyHostToDevice, s->stream));
        HIPX(hipStreamSynchronize(s->stream));
        s->nd = n;
    });
}

int lbm_sur_train(lbm_sur *s, int epochs, int batch, float lr, uint64_t seed, double *loss_per_epoch) {
    return guarded(s, [&] {
        if (epochs < 0 || batch <= 0 || !(lr > 0)) throw ArgError{LBM_ERR_ARG, "bad training parameters"};
        if (s->nd <= 0) throw ArgError{LBM_ERR_STATE, "no training data loaded"};
        const int B = std::min(batch, s->nd);
        ensure_batch(s, B);
        const int64_t ps = s->points(0, 1);
        std::vector<int> idx(s->nd);
        std::iota(idx.begin(), idx.end(), 0);
        std::mt19937_64 rng(seed);
        for (int e = 0; e < epochs; e++) {
            std::shuffle(idx.begin(), idx.end(), rng);
            double sum = 0.0;
            int nb = 0;
            for (int b0 = 0; b0 + B <= s->nd; b0 += B) {
                for (int j = 0; j < B; j++) {
                    const int k = idx[b0 + j];
                    HIPX(hipMemcpyAsync(s->act[X0] + (int64_t)j * ps * kInPad, s->din + (int64_t)k * ps * kInPad,
                                        sizeof(bf16) * ps * kInPad, hipMemcpyDeviceToDevice, s->stream));
                    HIPX(hipMemcpyAsync(s->tgt + (int64_t)j * ps * kOutUsed, s->dtg + (int64_t)k * ps * kOutUsed,
                                        sizeof(float) * ps * kOutUsed, hipMemcpyDeviceToDevice, s->stream));
                }
                forward(s, B);
                sum += loss_and_head_grad(s, B, true);
                nb++;
                backward(s, B);
                s->t_adam++;
                const double c1 = 1.0 - std::pow((double)kBeta1, (double)s->t_adam);
                const double c2 = 1.0 - std::pow((double)kBeta2, (double)s->t_adam);
                const float lr_t = (float)(lr * std::sqrt(c2) / c1), eps_t = (float)(kEps * std::sqrt(c2));
                hipLaunchKernelGGL(adam_kernel, dim3(grid_for(s->nparam)), dim3(256), 0, s->stream, s->w, s->grad,
                                   s->m, s->v, s->nparam, lr_t, kBeta1, kBeta2, eps_t);
                HIPX(hipGetLastError());
                cast_all(s);
            }
            if (loss_per_epoch) loss_per_epoch[e] = nb ? sum / nb : 0.0;
        }
        HIPX(hipStreamSynchronize(s->stream));
    });
}

int lbm_sur_predict(lbm_sur *s, int n, const float *inputs, float *fields) {
    return guarded(s, [&] {
        if (n <= 0 || !inputs || !fields) throw ArgError{LBM_ERR_ARG, "bad arguments"};
        const int B = std::max(1, std::min(n, std::max(s->cap, 8)));
        ensure_batch(s, B);
        const int64_t ps = s->points(0, 1);
        std::vector<float> o((size_t)(ps * 8 * B));
        for (int i0 = 0; i0 < n; i0 += B) {
            const int b = std::min(B, n - i0);
            upload_inputs(s, b, inputs + (int64_t)i0 * ps * kInUsed);
            forward(s, b);
            HIPX(hipMemcpyAsync(o.data(), s->out, sizeof(float) * ps * 8 * b, hipMemcpyDeviceToHost, s->stream));
            HIPX(hipStreamSynchronize(s->stream));
            for (int64_t p = 0; p < ps * b; p++)
                for (int c = 0; c < kOutUsed; c++) fields[((int64_t)i0 * ps + p) * kOutUsed + c] = o[(size_t)(p * 8 + c)];
        }
    });
}

int lbm_sur_warm_start(lbm_sur *s, lbm_handle *h, double Re, double u_inf) {
    return guarded(s, [&] {
        if (!h || !(Re > 0) || !std::isfinite(u_inf)) throw ArgError{LBM_ERR_ARG, "bad arguments"};
        int nx, ny, nz;
        lbm::handle_dims(h, &nx, &ny, &nz);
        if (nx != ny || ny != nz) throw ArgError{LBM_ERR_ARG, "warm start needs a cubic mesh"};
        const int S = s->S, N = nx;
        const int64_t ps = s->points(0, 1);
        std::vector<float> in((size_t)(ps * kInUsed));
        lbm_sur_encode(S, Re, N, in.data());
        // predict one sample (inlined to keep the error state on this object)
        ensure_batch(s, std::max(1, s->cap));
        upload_inputs(s, 1, in.data());
        forward(s, 1);
        HIPX(hipStreamSynchronize(s->stream));
        // trilinear interpolation onto the solver's nodes and the equilibrium
        // moments, on the device, straight into the solver's layout
        lbm::warm_moments(h, s->out, S, 8, u_inf);
    });
}

int lbm_solve_to_residual(lbm_handle *h, lbm_sur *s, double Re, double u_inf, double tol, int check,
                          int64_t max_steps, lbm_solve_info *info) {
    if (!h || !info || check <= 0 || max_steps < 0 || !(tol >= 0)) return LBM_ERR_ARG;
    std::memset(info, 0, sizeof *info);
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    int rc = LBM_OK;
    double ke = 0, r = INFINITY;
    int64_t steps = 0;
    auto ms_since = [&](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    if (s) {
        rc = lbm_sur_warm_start(s, h, Re, u_inf);
        if (rc != LBM_OK) return rc;
        rc = lbm_synchronize(h);
        if (rc != LBM_OK) return rc;
        info->ms_warm = ms_since(t0);
        info->warm = 1;
    }
    // reference field of the first check: the initial state
    rc = lbm_reduce(h, &ke, &r);
    if (rc != LBM_OK) return rc;
    bool first = true;
    while (steps < max_steps) {
        const int n = (int)std::min<int64_t>(check, max_steps - steps);
        rc = lbm_step(h, n);
        if (rc == LBM_OK) rc = lbm_reduce(h, &ke, &r);
        if (rc != LBM_OK) return rc;
        steps += n;
        if (first) {
            info->first_residual = r;
            first = false;
            if (s && (!std::isfinite(r) || !std::isfinite(ke) || r >= 1.0)) {
                // reject the warm start: back to rest
                int nx, ny, nz;
                lbm::handle_dims(h, &nx, &ny, &nz);
                const size_t nn = (size_t)nx * ny * nz;
                std::vector<float> zero(nn, 0.f);
                const float *mp[10];
                for (int k = 0; k < 10; k++) mp[k] = zero.data();
                rc = lbm_set_moments(h, mp, LBM_LAYOUT_GLOBAL);
                if (rc == LBM_OK) rc = lbm_reduce(h, &ke, &r);
                if (rc != LBM_OK) return rc;
                info->warm = 0;
                info->fell_back = 1;
                r = INFINITY;
                continue;
            }
        }
        if (r < tol) break;
    }
    rc = lbm_synchronize(h);
    if (rc != LBM_OK) return rc;
    info->steps = steps;
    info->residual = r;
    info->ms = ms_since(t0);
    return LBM_OK;
}

}  // extern "C"
