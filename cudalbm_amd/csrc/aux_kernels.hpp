// aux_kernels.hpp -- small kernels around the step: storage codec, step counter,
// FP64 reductions with a fixed reduction order.
#pragma once

#include "step_kernel.hpp"

namespace lbm {

// FP32 staging (internal node order) <-> FP16 encoded moment k (p's layout).
static __global__ void encode_f16_kernel(const StepParams p, int k, const float *__restrict__ src, float enc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < p.n; i += (int64_t)gridDim.x * blockDim.x)
        Store<_Float16>::store(p.mom[k], mom_off<_Float16>(p, i), _Float16(src[i] * enc));
}
static __global__ void decode_f16_kernel(const StepParams p, int k, float *__restrict__ dst, float dec) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < p.n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = Store<_Float16>::load(p.mom[k], mom_off<_Float16>(p, i), dec);
}

static __global__ void fill_kernel_f64(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
static __global__ void fill_kernel_f32(float *p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Instantaneous k and S (either output may be null) from the stored moments.
template <typename ST>
__global__ void fo_instant_kernel(const StepParams p, float *__restrict__ k, float *__restrict__ S) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < p.n; i += (int64_t)gridDim.x * blockDim.x) {
        float m[10];
        const unsigned off = mom_off<ST>(p, i);
#pragma unroll
        for (int c = 0; c < 10; c++) m[c] = Store<ST>::load(p.mom[c], off, p.dec[c]);
        float kk, s[6];
        fo_values(m, p.s_diag, p.s_off, kk, s);
        if (k) k[i] = kk;
        if (S) {
#pragma unroll
            for (int a = 0; a < 6; a++) S[a * p.n + i] = s[a];
        }
    }
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
// Block j of nb covers nodes j*NT + tid + k*nb*NT; part receives its 3 sums.
template <typename ST, int NT>
__device__ __forceinline__ void residual_partials(const StepParams &p, float *__restrict__ uprev,
                                                  double *__restrict__ part, int first, int j, int nb, double *sh) {
    double ke = 0.0, du = 0.0, uu = 0.0;
    for (int64_t i = j * (int64_t)NT + threadIdx.x; i < p.n; i += (int64_t)nb * NT) {
        // physical rho and u formed in FP64 from the stored values
        const unsigned off = mom_off<ST>(p, i);
        const double rho = 1.0 + (double)Store<ST>::load(p.mom[0], off, p.dec[0]);
        double u[3];
#pragma unroll
        for (int a = 0; a < 3; a++) u[a] = (double)Store<ST>::load(p.mom[1 + a], off, p.dec[1 + a]) / 3.0;
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
        part[0] = ke;
        part[1] = du;
        part[2] = uu;
    }
}

template <typename ST, int NT>
__global__ __launch_bounds__(NT) void residual_partials_kernel(const StepParams p, float *__restrict__ uprev,
                                                               double *__restrict__ part, int first) {
    __shared__ double sh[NT / 64];
    residual_partials<ST, NT>(p, uprev, part + 3 * blockIdx.x, first, blockIdx.x, gridDim.x, sh);
}

// Sums per-block partials in block order: block b of the grid reduces the
// nblocks partials starting at part + 3*nblocks*b into out[3b..3b+2].
template <int NT>
__global__ __launch_bounds__(NT) void residual_final_kernel(const double *__restrict__ part, int nblocks,
                                                            double *__restrict__ out) {
    __shared__ double sh[NT / 64];
    part += 3 * (int64_t)nblocks * blockIdx.x;
    out += 3 * blockIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nblocks; i += NT)
        for (int k = 0; k < 3; k++) a[k] += part[3 * i + k];
    for (int k = 0; k < 3; k++) {
        const double s = block_sum<NT>(a[k], sh);
        if (threadIdx.x == 0) out[k] = s;
    }
}

// Column-tiled node order of one footprint: node (x, y, z) at
// ((z nby + by) nbx + bx) tx ty + (y mod ty) tx + (x mod tx).
// mplane: elements from one plane block of an array to its next (tx ty for a
// plain array, 10 tx ty for the interleaved moment arrays).
struct ColumnTiling {
    int tx, ty, nbx, nby, mplane;
    __device__ __forceinline__ int64_t element(int x, int y, int z) const {
        const int bx = x / tx, by = y / ty;
        return (((int64_t)z * nby + by) * nbx + bx) * (int64_t)mplane + (int64_t)(y % ty) * tx + (x % tx);
    }
};

// One array of n nodes from the node order `from` into the order `to`
// (a permutation: dst[to(x)] = src[from(x)]); grid-stride over the destination nodes.
template <typename T>
__global__ void retile_kernel(const T *__restrict__ src, T *__restrict__ dst, int64_t n, ColumnTiling from,
                              ColumnTiling to) {
    const int64_t per = (int64_t)to.tx * to.ty, cols = (int64_t)to.nbx * to.nby;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t tile = i / per, r = i % per;
        const int z = (int)(tile / cols), c = (int)(tile % cols);
        const int by = c / to.nbx, bx = c % to.nbx;
        const int y = by * to.ty + (int)(r / to.tx), x = bx * to.tx + (int)(r % to.tx);
        dst[to.element(x, y, z)] = src[from.element(x, y, z)];
    }
}

// Warm start: a coarse S^3 field (channels u/U, v/U, w/U, p'/U^2 with
// p' = (rho - 1)/3, channel c of point q at f[q * cstride + c]) interpolated
// trilinearly onto the N^3 nodes (network node i at solver coordinate
// i (N-1)/(S-1)) and written as the equilibrium moments in the solver's
// layout: rho - 1 = 3 U^2 p', 3u, 4.5 u^2, 9 uv, ...; walls at rest, the lid
// row moving at U along +x.  One thread per node of the internal order.
template <typename ST>
__global__ void warm_moments_kernel(const StepParams p, ColumnTiling t, const float *__restrict__ f, int S,
                                    int cstride, double U) {
    const int N = p.nx;
    const int64_t per = (int64_t)t.tx * t.ty, cols = (int64_t)t.nbx * t.nby;
    auto axis = [&](int j, int &k, float &fr) {
        const double sc = N > 1 ? (double)j * (S - 1) / (N - 1) : 0.0;
        k = min(S - 2, (int)floor(sc));
        fr = (float)(sc - k);
    };
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < p.n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t tile = i / per, r = i % per;
        const int z = (int)(tile / cols), c = (int)(tile % cols);
        const int y = (c / t.nbx) * t.ty + (int)(r / t.tx), x = (c % t.nbx) * t.tx + (int)(r % t.tx);
        int kx, ky, kz;
        float fx, fy, fz;
        axis(x, kx, fx);
        axis(y, ky, fy);
        axis(z, kz, fz);
        float val[4];
        for (int ch = 0; ch < 4; ch++) {
            float acc = 0.f;
            for (int d = 0; d < 8; d++) {
                const int ix = kx + (d & 1), iy = ky + ((d >> 1) & 1), iz = kz + (d >> 2);
                const float wx = (d & 1) ? fx : 1.f - fx;
                const float wy = ((d >> 1) & 1) ? fy : 1.f - fy;
                const float wz = (d >> 2) ? fz : 1.f - fz;
                acc += wx * wy * wz * f[(((int64_t)iz * S + iy) * S + ix) * cstride + ch];
            }
            val[ch] = acc;
        }
        double u = U * val[0], v = U * val[1], w = U * val[2];
        const double rm1 = 3.0 * U * U * val[3];
        if (y == N - 1) {
            u = U;
            v = w = 0.0;
        } else if (x == 0 || x == N - 1 || y == 0 || z == 0 || z == N - 1) {
            u = v = w = 0.0;
        }
        const float m[10] = {(float)rm1, (float)(3.0 * u), (float)(3.0 * v), (float)(3.0 * w), (float)(4.5 * u * u),
                             (float)(9.0 * u * v), (float)(9.0 * u * w), (float)(4.5 * v * v), (float)(9.0 * v * w),
                             (float)(4.5 * w * w)};
        const unsigned off = mom_off<ST>(p, i);
        for (int k = 0; k < 10; k++) {
            if constexpr (sizeof(ST) == 4)
                Store<ST>::store(p.mom[k], off, m[k]);
            else
                Store<ST>::store(p.mom[k], off, _Float16(m[k] * p.enc[k]));
        }
    }
}

// Kinetic energy and residual r = sqrt(sum |u - u_prev|^2 / sum |u|^2) from the
// three sums of residual_final_kernel, on the device (one thread).
static __global__ void residual_result_kernel(const double *__restrict__ sums, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double du = sums[1], uu = sums[2];
    out[0] = sums[0];
    out[1] = uu > 0 ? sqrt(du / uu) : (du > 0 ? INFINITY : 0.0);
}

}  // namespace lbm
