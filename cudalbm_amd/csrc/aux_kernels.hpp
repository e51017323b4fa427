// aux_kernels.hpp -- small kernels around the step: storage codec, step counter,
// FP64 reductions with a fixed reduction order.
#pragma once

#include "step_kernel.hpp"

namespace lbm {

// FP32 staging <-> FP16 encoded moments.
__global__ void encode_f16_kernel(const float *__restrict__ src, _Float16 *__restrict__ dst, int64_t n,
                                  float enc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = _Float16(src[i] * enc);
}
__global__ void decode_f16_kernel(const _Float16 *__restrict__ src, float *__restrict__ dst, int64_t n,
                                  float dec) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = float(src[i]) * dec;
}

// Advances the absolute step counter read by the fused time averages.
__global__ void tick_kernel(int64_t *t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) t[0] += 1;
}

__global__ void fill_kernel_f64(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_kernel_f32(float *p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Wave64 sum by butterfly shuffles, then across the waves of the block via LDS.
// The order of additions depends only on the thread index: deterministic.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; i++) r += sh[i];
    __syncthreads();
    return r;
}

// Per-block partial sums of kinetic energy, |u - u_prev|^2 and |u|^2 (physical u),
// then u_prev <- u.  Blocks stride over the node range in a fixed pattern.
template <typename ST, int NT>
__global__ __launch_bounds__(NT) void residual_partials_kernel(const StepParams p, float *__restrict__ uprev,
                                                               double *__restrict__ part, int first) {
    __shared__ double sh[NT / 64];
    double ke = 0.0, du = 0.0, uu = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < p.n; i += (int64_t)gridDim.x * NT) {
        // physical rho and u formed in FP64 from the stored values
        const double rho = 1.0 + (double)Store<ST>::load(p.mom[0], (int)i, p.dec[0]);
        double u[3];
#pragma unroll
        for (int a = 0; a < 3; a++) u[a] = (double)Store<ST>::load(p.mom[1 + a], (int)i, p.dec[1 + a]) / 3.0;
        double uu_i = 0.0, du_i = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            uu_i += u[a] * u[a];
            const double d = u[a] - (double)uprev[a * p.n + i];
            du_i += d * d;
            uprev[a * p.n + i] = (float)u[a];
        }
        ke += 0.5 * rho * uu_i;
        uu += uu_i;
        du += first ? uu_i : du_i;
    }
    ke = block_sum<NT>(ke, sh);
    du = block_sum<NT>(du, sh);
    uu = block_sum<NT>(uu, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = ke;
        part[3 * blockIdx.x + 1] = du;
        part[3 * blockIdx.x + 2] = uu;
    }
}

// Sums the per-block partials in block order (one block of NT threads).
template <int NT>
__global__ __launch_bounds__(NT) void residual_final_kernel(const double *__restrict__ part, int nblocks,
                                                            double *__restrict__ out) {
    __shared__ double sh[NT / 64];
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblocks; i += NT)
        for (int k = 0; k < 3; k++) a[k] += part[3 * i + k];
    for (int k = 0; k < 3; k++) {
        const double s = block_sum<NT>(a[k], sh);
        if (threadIdx.x == 0) out[k] = s;
    }
}

}  // namespace lbm
