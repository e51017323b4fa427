// lattice.hpp -- compile-time velocity sets and pointwise moment algebra shared by
// every kernel of the solver (gfx950).
//
// Populations are carried as deviations from the rest weights, g_q = f_q - w_q,
// so that the moment sums never cancel against the O(1) rest state in FP32:
//   rho - 1 = sum g_q,   rho u_a = sum g_q c_a,
//   Pi_ab = sum g_q c_a c_b + cs2 delta_ab.
// A population whose source lies outside the domain is carried as g = 0, which
// makes the reference's masked boundary sums (velocitySet.cuh:337-396) plain
// sums plus a per-node-type constant (the weights of the known populations).
//
// Velocity-set tables follow D3Q19.cuh:98-288 and D3Q27.cuh:98-225; the
// population index order and the face slot order (ascending q among the
// populations crossing a face) are the reference's, so that face buffers and
// probes speak the same population numbering.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int kC[27][3] = {
    {0, 0, 0},   {1, 0, 0},    {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
    {0, 0, -1},  {1, 1, 0},    {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},
    {0, -1, -1}, {1, -1, 0},   {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
    {0, -1, 1},  {1, 1, 1},    {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
    {-1, 1, -1}, {-1, 1, 1},   {1, -1, -1}};

__host__ __device__ constexpr int cxq(int q) { return kC[q][0]; }
__host__ __device__ constexpr int cyq(int q) { return kC[q][1]; }
__host__ __device__ constexpr int czq(int q) { return kC[q][2]; }
__host__ __device__ constexpr int cq(int q, int a) { return kC[q][a]; }
__host__ __device__ constexpr int absi(int a) { return a < 0 ? -a : a; }

template <int Q>
__host__ __device__ constexpr double wq(int q) {
    return Q == 19 ? (absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 0   ? 1.0 / 3.0
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 1 ? 1.0 / 18.0
                                                                        : 1.0 / 36.0)
                   : (absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 0   ? 8.0 / 27.0
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 1 ? 2.0 / 27.0
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 2 ? 1.0 / 54.0
                                                                        : 1.0 / 216.0);
}

template <int Q>
struct VS {
    static constexpr int QF = Q == 19 ? 5 : 9;  // populations crossing one face
    static constexpr int NSLOT = Q - 3;         // LDS slots: every q but 0, 5, 6
};

// LDS slot of a population that moves in-plane (q not in {0,5,6}).
__host__ __device__ constexpr int lds_slot(int q) { return q <= 4 ? q - 1 : q - 3; }

// k-th population (ascending q) with c[axis] == sign; the reference face-slot
// order (halo.cuh:813-904).
template <int Q>
__host__ __device__ constexpr int face_q(int axis, int sign, int k) {
    int n = 0;
    for (int q = 1; q < Q; q++)
        if (cq(q, axis) == sign) {
            if (n == k) return q;
            n++;
        }
    return -1;
}
template <int Q>
__host__ __device__ constexpr int face_k(int axis, int q) {
    int n = 0;
    for (int p = 1; p < Q; p++)
        if (cq(p, axis) == cq(q, axis)) {
            if (p == q) return n;
            n++;
        }
    return -1;
}

// ---------------------------------------------------------------- constants
constexpr float kCs2 = 1.0f / 3.0f;

// Compile-time loop.
template <int B, int E, typename F>
__device__ __forceinline__ void unroll(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        unroll<B + 1, E>(f);
    }
}

// Moment vector m: [0] rho-1, [1..3] 3u, [4] mxx, [5] mxy, [6] mxz, [7] myy,
// [8] myz, [9] mzz (scaled, as stored: velocitySet.cuh:130-142).
struct Rec {
    float rpw_base;  // rho - 1
    float rho;       // rho
    float A;         // -cs2 (mxx + myy + mzz)
};

__device__ __forceinline__ Rec rec_setup(const float (&m)[10]) {
    Rec r;
    r.rpw_base = m[0];
    r.rho = 1.0f + m[0];
    r.A = -kCs2 * (m[4] + m[7] + m[9]);
    return r;
}

// g_q = f_q - w_q with f_q = rho w_q [1 + B_q] (D3Q19.cuh:346-374): g_q = w_q (rho' + rho B_q).
template <int Q, int q>
__device__ __forceinline__ float rec_dev(const Rec &r, const float (&m)[10]) {
    constexpr int x = cxq(q), y = cyq(q), z = czq(q);
    float B = r.A;
    if constexpr (x != 0) B += float(x) * m[1];
    if constexpr (y != 0) B += float(y) * m[2];
    if constexpr (z != 0) B += float(z) * m[3];
    if constexpr (x != 0) B += m[4];
    if constexpr (y != 0) B += m[7];
    if constexpr (z != 0) B += m[9];
    if constexpr (x * y != 0) B += float(x * y) * m[5];
    if constexpr (x * z != 0) B += float(x * z) * m[6];
    if constexpr (y * z != 0) B += float(y * z) * m[8];
    constexpr float w = float(wq<Q>(q));
    return w * fmaf(r.rho, B, r.rpw_base);
}

// Raw second-order sums of the deviation populations.
struct Sums {
    float s0, sx, sy, sz, sxx, sxy, sxz, syy, syz, szz;
};

template <int Q>
__device__ __forceinline__ Sums pop_sums(const float (&g)[27]) {
    Sums s{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int x = cxq(q), y = cyq(q), z = czq(q);
        const float v = g[q];
        s.s0 += v;
        if constexpr (x != 0) s.sx += float(x) * v;
        if constexpr (y != 0) s.sy += float(y) * v;
        if constexpr (z != 0) s.sz += float(z) * v;
        if constexpr (x != 0) s.sxx += v;
        if constexpr (y != 0) s.syy += v;
        if constexpr (z != 0) s.szz += v;
        if constexpr (x * y != 0) s.sxy += float(x * y) * v;
        if constexpr (x * z != 0) s.sxz += float(x * z) * v;
        if constexpr (y * z != 0) s.syz += float(y * z) * v;
    });
    return s;
}

// Interior moments (velocitySet.cuh:398-417) followed by scale (velocitySet.cuh:130-142).
__device__ __forceinline__ void interior_moments(const Sums &s, float (&m)[10]) {
    const float rho = 1.0f + s.s0;
    const float inv = 1.0f / rho;
    m[0] = s.s0;
    m[1] = 3.0f * s.sx * inv;
    m[2] = 3.0f * s.sy * inv;
    m[3] = 3.0f * s.sz * inv;
    m[4] = 4.5f * (fmaf(-kCs2, s.s0, s.sxx) * inv);
    m[5] = 9.0f * s.sxy * inv;
    m[6] = 9.0f * s.sxz * inv;
    m[7] = 4.5f * (fmaf(-kCs2, s.s0, s.syy) * inv);
    m[8] = 9.0f * s.syz * inv;
    m[9] = 4.5f * (fmaf(-kCs2, s.s0, s.szz) * inv);
}

// Boundary moments: boundaryConditions.cuh:99-554, in scaled form.  `bits` is the
// normalVector bitmask (W=1, E=2, S=4, N=8, B=16, F=32).
template <int Q>
__device__ __forceinline__ void boundary_moments(const Sums &s, int bits, float omega, float u_lid,
                                                 float (&m)[10]) {
    // Known-population weight sums for this node type (incomingSwitch,
    // velocitySet.cuh:204-220): rho_I = W + sum g, Sab_I = Wab + sum g c_a c_b.
    float W = 0.f, Wxy = 0.f, Wxz = 0.f, Wyz = 0.f;
    unroll<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int x = cxq(q), y = cyq(q), z = czq(q);
        bool unknown = false;
        if constexpr (x > 0) unknown |= (bits & 1) != 0;
        if constexpr (x < 0) unknown |= (bits & 2) != 0;
        if constexpr (y > 0) unknown |= (bits & 4) != 0;
        if constexpr (y < 0) unknown |= (bits & 8) != 0;
        if constexpr (z > 0) unknown |= (bits & 16) != 0;
        if constexpr (z < 0) unknown |= (bits & 32) != 0;
        constexpr float w = float(wq<Q>(q));
        const float k = unknown ? 0.f : w;
        W += k;
        if constexpr (x * y != 0) Wxy += float(x * y) * k;
        if constexpr (x * z != 0) Wxz += float(x * z) * k;
        if constexpr (y * z != 0) Wyz += float(y * z) * k;
    });
    const float rho_I = W + s.s0;
    const float inv_I = 1.0f / rho_I;
    const float mxy_I = (Wxy + s.sxy) * inv_I;
    const float mxz_I = (Wxz + s.sxz) * inv_I;
    const float myz_I = (Wyz + s.syz) * inv_I;

    const bool north = (bits & 8) != 0;
    const int nb = __popc(bits & 0x3F);
    const float ux = north ? u_lid / float(nb) : 0.f;  // U<X> (boundaryConditions.cuh:543-554)
    const float w = omega, U = u_lid;
    float rho = 1.f, mxy = 0.f, mxz = 0.f, myz = 0.f;
    auto edge = [&](float sign, float mI, float &mab) {
        // rho = 36 (rho_I - mI rho_I + mI rho_I w)/(24 + w) for sign +1, and
        // -36(-rho_I - mI rho_I + mI rho_I w)/(24 + w) for sign -1.
        const float a = mI * rho_I;
        rho = (sign > 0.f) ? 36.f * (rho_I - a + a * w) / (24.f + w)
                           : -36.f * (-rho_I - a + a * w) / (24.f + w);
        mab = (36.f * a - sign * rho) / (9.f * rho);
    };
    switch (bits & 0x3F) {
    case 21: case 37: case 22: case 38:  // SWB SWF SEB SEF
        rho = (Q == 19) ? 12.f * rho_I / 7.f : 216.f * rho_I / 125.f;
        break;
    case 5: edge(+1.f, mxy_I, mxy); break;   // SW
    case 6: edge(-1.f, mxy_I, mxy); break;   // SE
    case 17: edge(+1.f, mxz_I, mxz); break;  // WB
    case 33: edge(-1.f, mxz_I, mxz); break;  // WF
    case 18: edge(-1.f, mxz_I, mxz); break;  // EB
    case 34: edge(+1.f, mxz_I, mxz); break;  // EF
    case 20: edge(+1.f, myz_I, myz); break;  // SB
    case 36: edge(-1.f, myz_I, myz); break;  // SF
    case 1: case 2:  // W, E
        rho = 6.f * rho_I / 5.f;
        mxy = 2.f * mxy_I * rho_I / rho;
        mxz = 2.f * mxz_I * rho_I / rho;
        break;
    case 4:  // S
        rho = 6.f * rho_I / 5.f;
        mxy = 2.f * mxy_I * rho_I / rho;
        myz = 2.f * myz_I * rho_I / rho;
        break;
    case 16: case 32:  // B, F
        rho = 6.f * rho_I / 5.f;
        mxz = 2.f * mxz_I * rho_I / rho;
        myz = 2.f * myz_I * rho_I / rho;
        break;
    case 8:  // N (moving lid)
        rho = 6.f * rho_I / 5.f;
        mxy = (6.f * mxy_I * rho_I - U * rho) / (3.f * rho);
        myz = 2.f * myz_I * rho_I / rho;
        break;
    case 25: case 41:  // NWB NWF
        rho = -24.f * rho_I / (-14.f - 8.f * U + 9.f * U * U);
        break;
    case 26: case 42:  // NEB NEF
        rho = -24.f * rho_I / (-14.f + 8.f * U + 9.f * U * U);
        break;
    case 24: {  // NB
        const float a = myz_I * rho_I;
        rho = 72.f * (-rho_I - a + a * w) / (-48.f - 2.f * w + 3.f * U * U * w);
        myz = (72.f * a + 2.f * rho - 3.f * U * U * rho) / (18.f * rho);
    } break;
    case 40: {  // NF
        const float a = myz_I * rho_I;
        rho = -72.f * (rho_I - a + a * w) / (-48.f - 2.f * w + 3.f * U * U * w);
        myz = (72.f * a - 2.f * rho + 3.f * U * U * rho) / (18.f * rho);
    } break;
    case 10: {  // NE
        const float a = mxy_I * rho_I;
        rho = 36.f * (rho_I - a + a * w) /
              (24.f - 18.f * U - 18.f * U * U + w + 3.f * U * w + 3.f * U * U * w);
        mxy = (36.f * a - rho - 3.f * U * rho - 3.f * U * U * rho) / (9.f * rho);
    } break;
    case 9: {  // NW
        const float a = mxy_I * rho_I;
        rho = -36.f * (-rho_I - a + a * w) /
              (24.f + 18.f * U - 18.f * U * U + w - 3.f * U * w + 3.f * U * U * w);
        mxy = (36.f * a + rho - 3.f * U * rho + 3.f * U * U * rho) / (9.f * rho);
    } break;
    default: break;
    }
    // unscaled -> scaled (velocitySet.cuh:130-142); m_xx = u_x^2, m_yy = m_zz = 0
    m[0] = rho - 1.0f;
    m[1] = 3.f * ux;
    m[2] = 0.f;
    m[3] = 0.f;
    m[4] = 4.5f * ux * ux;
    m[5] = 9.f * mxy;
    m[6] = 9.f * mxz;
    m[7] = 0.f;
    m[8] = 9.f * myz;
    m[9] = 0.f;
}

// Second-order regularised collision in scaled moment space (secondOrder.cuh:93-105).
__device__ __forceinline__ void collide(float (&m)[10], float t_omega, float omega_h, float omega) {
    m[4] = fmaf(t_omega, m[4], omega_h * m[1] * m[1]);
    m[7] = fmaf(t_omega, m[7], omega_h * m[2] * m[2]);
    m[9] = fmaf(t_omega, m[9], omega_h * m[3] * m[3]);
    m[5] = fmaf(t_omega, m[5], omega * m[1] * m[2]);
    m[6] = fmaf(t_omega, m[6], omega * m[1] * m[3]);
    m[8] = fmaf(t_omega, m[8], omega * m[2] * m[3]);
}

// Node-type bitmask (normalVector.cuh:321-334).
__host__ __device__ __forceinline__ int node_bits(int x, int y, int z, int nx, int ny, int nz) {
    const int b = (x == 0) | ((x == nx - 1) << 1) | ((y == 0) << 2) | ((y == ny - 1) << 3) |
                  ((z == 0) << 4) | ((z == nz - 1) << 5);
    return b | ((b != 0) << 6);
}

}  // namespace lbm
