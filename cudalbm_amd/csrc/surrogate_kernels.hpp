// surrogate_kernels.hpp -- BF16 MFMA kernels of the 3D U-Net surrogate.
//
// Tensors are channels-last, [P][C] with point p = ((b*S + z)*S + y)*S + x on
// an S^3 grid (batch b).  A 3x3x3 convolution with zero padding is an
// implicit GEMM over K = taps*C (tap-major, channels fastest):
//   forward   Y[p][co]  = sum_{t,ci} X[p+o(t)][ci] W[co][t][ci]        (M=P,    N=Cout, K=27 Cin)
//   input     dX[p][ci] = sum_{t,co} G[p-o(t)][co] Wt[ci][t][co]       (M=P,    N=Cin,  K=27 Cout)
//   weight    dW[co][t][ci] = sum_p G[p][co] X[p+o(t)][ci]             (M=Cout, N=27 Cin, K=P)
// with G the gradient at the pre-activation and Wt the transposed weights.
// Each wave computes a 16 x (16*NB) block of the GEMM with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation).  Operand lane layout
// (checked on the GPU): lane l holds A row l%16, B column l%16, both for the
// eight K indices 8*(l/16)..8*(l/16)+7; the result lane l holds column l%16,
// rows 4*(l/16)..4*(l/16)+3.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sur {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { FWD = 0, DGRAD = 1, WGRAD = 2 };

struct ConvArgs {
    const bf16 *x;      // FWD, WGRAD: input activations [P][Cin]
    const bf16 *g;      // DGRAD, WGRAD: gradient at the pre-activation [P][Cout]
    const bf16 *w;      // FWD: [Cout][T*Cin]; DGRAD: transposed [Cin][T*Cout]
    const float *bias;  // FWD: [Cout]
    bf16 *y;            // FWD with ReLU: output [P][Cout]
    float *yf;          // FWD without activation: output [P][Cout]; DGRAD: dX [P][Cin]
    float *part;        // WGRAD: split-K partials [chunk][Cout][T*Cin]
    int S, P, Cin, Cout, T, relu, kchunk;
};

// Offset of tap t (T = 27: 3x3x3 stencil, T = 1: centre only) and whether the
// neighbour of (x,y,z) exists.
__device__ __forceinline__ bool tap_src(int t, int T, int S, int x, int y, int z, int sign, int *off) {
    if (T == 1) {
        *off = 0;
        return true;
    }
    const int dz = sign * (t / 9 - 1), dy = sign * ((t / 3) % 3 - 1), dx = sign * (t % 3 - 1);
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    *off = (dz * S + dy) * S + dx;
    return (unsigned)xx < (unsigned)S && (unsigned)yy < (unsigned)S && (unsigned)zz < (unsigned)S;
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = (bf16)0.f;
    return v;
}

// FWD and DGRAD: M = points.  Block = 4 waves on consecutive 16-point rows,
// blockIdx.y selects the 16*NB-wide column block.
template <int MODE, int NB>
__global__ __launch_bounds__(256) void conv_points_kernel(const ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const int m0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (m0 >= a.P) return;
    const int n0 = blockIdx.y * 16 * NB;
    // FWD: A = X (C = Cin), B = W (N = Cout);  DGRAD: A = G (C = Cout), B = Wt (N = Cin)
    const int C = MODE == FWD ? a.Cin : a.Cout;
    const int N = MODE == FWD ? a.Cout : a.Cin;
    const int K = a.T * C;
    const bf16 *A = MODE == FWD ? a.x : a.g;
    const int sign = MODE == FWD ? 1 : -1;
    const int pa = m0 + (lane & 15);
    const int S = a.S;
    const int ax = pa % S, ay = (pa / S) % S, az = (pa / (S * S)) % S;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int kg = k0 + 8 * (lane >> 4);
        bf16x8 av = zero8();
        if (kg < K && pa < a.P) {
            const int t = kg / C, c = kg - t * C;
            int off;
            if (tap_src(t, a.T, S, ax, ay, az, sign, &off))
                av = *reinterpret_cast<const bf16x8 *>(A + (int64_t)(pa + off) * C + c);
        }
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int n = n0 + 16 * nb + (lane & 15);
            bf16x8 bv = zero8();
            if (kg < K && n < N) bv = *reinterpret_cast<const bf16x8 *>(a.w + (int64_t)n * K + kg);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int n = n0 + 16 * nb + (lane & 15);
        if (n >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int p = m0 + 4 * (lane >> 4) + r;
            if (p >= a.P) continue;
            float v = acc[nb][r];
            if (MODE == FWD) {
                v += a.bias[n];
                if (a.relu)
                    a.y[(int64_t)p * N + n] = (bf16)fmaxf(v, 0.f);
                else
                    a.yf[(int64_t)p * N + n] = v;
            } else {
                a.yf[(int64_t)p * N + n] = v;
            }
        }
    }
}

// WGRAD: M = Cout, N = T*Cin, K = the points of chunk blockIdx.y.  One wave
// per block; blockIdx.x enumerates (m block, n block).
template <int NB>
__global__ __launch_bounds__(64) void conv_wgrad_kernel(const ConvArgs a) {
    const int lane = threadIdx.x;
    const int N = a.T * a.Cin;
    const int nblk = (N + 16 * NB - 1) / (16 * NB);
    const int m0 = (blockIdx.x / nblk) * 16, n0 = (blockIdx.x % nblk) * 16 * NB;
    const int p_begin = blockIdx.y * a.kchunk;
    const int p_end = min(a.P, p_begin + a.kchunk);
    const int S = a.S;
    const int co = m0 + (lane & 15);
    int tn[NB], cn[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int n = n0 + 16 * nb + (lane & 15);
        tn[nb] = n < N ? n / a.Cin : -1;
        cn[nb] = n < N ? n - tn[nb] * a.Cin : 0;
    }
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = p_begin; k0 < p_end; k0 += 32) {
        const int kp = k0 + 8 * (lane >> 4);
        bf16x8 av = zero8();
        if (co < a.Cout) {
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (kp + e < p_end) av[e] = a.g[(int64_t)(kp + e) * a.Cout + co];
        }
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            bf16x8 bv = zero8();
            if (tn[nb] >= 0) {
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int p = kp + e;
                    if (p >= p_end) continue;
                    int off;
                    if (tap_src(tn[nb], a.T, S, p % S, (p / S) % S, (p / (S * S)) % S, 1, &off))
                        bv[e] = a.x[(int64_t)(p + off) * a.Cin + cn[nb]];
                }
            }
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[nb], 0, 0, 0);
        }
    }
    // result lane l: column n0+16nb+l%16, rows m0 + 4(l/16) + r
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        const int n = n0 + 16 * nb + (lane & 15);
        if (n >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + 4 * (lane >> 4) + r;
            if (m < a.Cout) a.part[((int64_t)blockIdx.y * a.Cout + m) * N + n] = acc[nb][r];
        }
    }
}

// dW[i] = sum over chunks of part (fixed order).
__global__ void sum_chunks_kernel(const float *__restrict__ part, int nchunk, int64_t len, float *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nchunk; c++) s += part[c * len + i];
        out[i] = s;
    }
}

// db[c] = sum_p G[p][c]: block c, FP64 accumulation, fixed order.
__global__ __launch_bounds__(256) void bias_grad_kernel(const bf16 *__restrict__ g, int P, int C, float *__restrict__ db) {
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int p = threadIdx.x; p < P; p += 256) s += (double)(float)g[(int64_t)p * C + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = (float)sh[0];
}

// G = bf16(dY * (Y > 0)): gradient at the pre-activation of a ReLU layer.
__global__ void relu_grad_kernel(const float *__restrict__ dy, const bf16 *__restrict__ y, int64_t n,
                                 bf16 *__restrict__ g) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        g[i] = (bf16)((float)y[i] > 0.f ? dy[i] : 0.f);
}

// 2x2x2 average pooling S -> S/2 (bf16 in and out, FP32 arithmetic).
__global__ void pool_fwd_kernel(const bf16 *__restrict__ in, int S, int B, int C, bf16 *__restrict__ out) {
    const int h = S / 2;
    const int64_t n = (int64_t)B * h * h * h * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        int64_t q = i / C;
        const int x = q % h, y = (q / h) % h, z = (q / ((int64_t)h * h)) % h, b = q / ((int64_t)h * h * h);
        float s = 0.f;
        for (int d = 0; d < 8; d++) {
            const int64_t p = (((int64_t)b * S + 2 * z + (d >> 2)) * S + 2 * y + ((d >> 1) & 1)) * S + 2 * x + (d & 1);
            s += (float)in[p * C + c];
        }
        out[i] = (bf16)(s * 0.125f);
    }
}

// dIn[child] += dOut[parent] / 8 (FP32).
__global__ void pool_bwd_add_kernel(const float *__restrict__ dout, int S, int B, int C, float *__restrict__ din) {
    const int64_t n = (int64_t)B * S * S * S * C;
    const int h = S / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        const int64_t p = i / C;
        const int x = p % S, y = (p / S) % S, z = (p / ((int64_t)S * S)) % S, b = p / ((int64_t)S * S * S);
        const int64_t q = (((int64_t)b * h + z / 2) * h + y / 2) * h + x / 2;
        din[i] += dout[q * C + c] * 0.125f;
    }
}

// out[p] = [ low[parent(p)] (Ca channels, nearest 2x upsampling) | skip[p] (Cb channels) ]
__global__ void upcat_fwd_kernel(const bf16 *__restrict__ low, const bf16 *__restrict__ skip, int S, int B, int Ca,
                                 int Cb, bf16 *__restrict__ out) {
    const int C = Ca + Cb, h = S / 2;
    const int64_t n = (int64_t)B * S * S * S * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = i % C;
        const int64_t p = i / C;
        if (c < Ca) {
            const int x = p % S, y = (p / S) % S, z = (p / ((int64_t)S * S)) % S, b = p / ((int64_t)S * S * S);
            const int64_t q = (((int64_t)b * h + z / 2) * h + y / 2) * h + x / 2;
            out[i] = low[q * Ca + c];
        } else {
            out[i] = skip[p * Cb + (c - Ca)];
        }
    }
}

// dLow[q] = sum of the 8 children's first Ca channels; dSkip[p] = last Cb channels.
__global__ void upcat_bwd_kernel(const float *__restrict__ dcat, int S, int B, int Ca, int Cb,
                                 float *__restrict__ dlow, float *__restrict__ dskip) {
    const int C = Ca + Cb, h = S / 2;
    const int64_t nl = (int64_t)B * h * h * h * Ca, ns = (int64_t)B * S * S * S * Cb;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl + ns; i += (int64_t)gridDim.x * blockDim.x) {
        if (i < nl) {
            const int c = i % Ca;
            const int64_t q = i / Ca;
            const int x = q % h, y = (q / h) % h, z = (q / ((int64_t)h * h)) % h, b = q / ((int64_t)h * h * h);
            float s = 0.f;
            for (int d = 0; d < 8; d++) {
                const int64_t p = (((int64_t)b * S + 2 * z + (d >> 2)) * S + 2 * y + ((d >> 1) & 1)) * S + 2 * x + (d & 1);
                s += dcat[p * C + c];
            }
            dlow[i] = s;
        } else {
            const int64_t j = i - nl;
            const int c = j % Cb;
            const int64_t p = j / Cb;
            dskip[j] = dcat[p * C + Ca + c];
        }
    }
}

// Mean-squared error over the first 4 of the head's 8 output channels (the
// last 4 only pad the channel count to a multiple of 8 and carry no loss):
// G = bf16(2 (out - tgt) / (4P)) on channels 0-3 and 0 on 4-7; per-block FP64
// partial sums of (out - tgt)^2.  out [P][8], tgt [P][4].
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ out, const float *__restrict__ tgt, int64_t P,
                                                  bf16 *__restrict__ g, double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    const float sc = 2.0f / (float)(4 * P);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 8 * P; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = i & 7;
        const int64_t p = i >> 3;
        float d = 0.f;
        if (c < 4) {
            d = out[i] - tgt[p * 4 + c];
            s += (double)d * d;
        }
        if (g) g[i] = (bf16)(sc * d);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on the FP32 master parameters (bias-corrected step sizes passed in).
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int64_t n, float lr_t, float b1, float b2, float eps_t) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
    }
}

// BF16 copies of one convolution's weights: wb[co][t][ci] and wt[ci][t][co].
__global__ void cast_weights_kernel(const float *__restrict__ w, int Cout, int T, int Cin, bf16 *__restrict__ wb,
                                    bf16 *__restrict__ wt) {
    const int64_t n = (int64_t)Cout * T * Cin;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int ci = i % Cin;
        const int t = (i / Cin) % T;
        const int co = i / ((int64_t)Cin * T);
        const bf16 b = (bf16)w[i];
        wb[i] = b;
        wt[((int64_t)ci * T + t) * Cout + co] = b;
    }
}

// FP32 [P][c_src] (first c_src channels) -> BF16 [P][c_dst] with zero padding.
__global__ void pack_input_kernel(const float *__restrict__ src, int64_t P, int c_src, int c_dst,
                                  bf16 *__restrict__ dst) {
    const int64_t n = P * c_dst;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = i % c_dst;
        const int64_t p = i / c_dst;
        dst[i] = (bf16)(c < c_src ? src[p * c_src + c] : 0.f);
    }
}

}  // namespace sur
