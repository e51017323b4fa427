// lattice.hpp -- compile-time velocity sets and pointwise moment algebra shared by
// every kernel of the solver (gfx950).
//
// Populations are carried as scaled deviations from the rest weights,
// H_q = 3 D (f_q - w_q), so that the moment sums never cancel against the O(1)
// rest state in FP32 and every constant is exact:
//   3 D (rho - 1) = sum H_q,   3 D rho u_a = sum H_q c_a,
//   3 D (Pi_ab - cs2 delta_ab) = sum H_q c_a c_b.
// A population whose source lies outside the domain is carried as H = 0, which
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
    static constexpr int NU = Q == 19 ? 4 : 8;  // up-going pops that also move in-plane
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
// Integer weights: w_q = Wint_q / D with D = 36 (D3Q19) or 216 (D3Q27).
template <int Q>
__host__ __device__ constexpr int wint(int q) {
    return Q == 19 ? (absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 0   ? 12
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 1 ? 2
                                                                        : 1)
                   : (absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 0   ? 64
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 1 ? 16
                      : absi(cxq(q)) + absi(cyq(q)) + absi(czq(q)) == 2 ? 4
                                                                        : 1);
}
template <int Q>
__host__ __device__ constexpr float d3() {  // 3 D
    return Q == 19 ? 108.f : 648.f;
}

// Compile-time loop.
template <int B, int E, typename F>
__device__ __forceinline__ void unroll(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        unroll<B + 1, E>(f);
    }
}

// Populations travel as H_q = 3 D (f_q - w_q).  Every constant of the
// reconstruct / moment round trip is then an exact small integer or dyadic, so
// FP32 rounding is unbiased: roundings of 1/18, 1/36 or 1/3 would otherwise
// scale the state by the same factor on every step.
//
// Moment vector m: [0] rho-1, [1..3] 3u, [4] mxx, [5] mxy, [6] mxz, [7] myy,
// [8] myz, [9] mzz (scaled, as stored: velocitySet.cuh:130-142).
struct Rec {
    float a;     // 3 rho' - rho (mxx + myy + mzz)
    float M[10]; // 3 rho m_k, k = 1..9 (M[0] unused)
};

__device__ __forceinline__ Rec rec_setup(const float (&m)[10]) {
    Rec r;
    const float rho = 1.0f + m[0];
    r.a = fmaf(rho, -(m[4] + m[7] + m[9]), 3.0f * m[0]);
    const float t = 3.0f * rho;
#pragma unroll
    for (int k = 1; k < 10; k++) r.M[k] = t * m[k];
    r.M[0] = 0.f;
    return r;
}

// f_q = rho w_q [1 - cs2(mxx+myy+mzz) + c.m + ...] (D3Q19.cuh:346-374), so
// H_q = Wint_q (a + sum_a c_a M_a + sum_a c_a^2 M_aa + sum_{a<b} c_a c_b M_ab).
// The even part (diagonal, then off-diagonal terms) and the odd part are formed
// in one fixed order with a canonical sign, so a pair q / opposite(q) and all
// populations of one plane share them (common subexpressions).
template <int Q, int q>
__device__ __forceinline__ float rec_dev(const Rec &r) {
    constexpr int x = cxq(q), y = cyq(q), z = czq(q);
    float e = r.a;
    if constexpr (x != 0) e += r.M[4];
    if constexpr (y != 0) e += r.M[7];
    if constexpr (z != 0) e += r.M[9];
    // off-diagonal part, canonical sign: first nonzero product positive
    constexpr int pxy = x * y, pxz = x * z, pyz = y * z;
    constexpr int s_off = pxy != 0 ? pxy : (pxz != 0 ? pxz : (pyz != 0 ? pyz : 1));
    if constexpr (pxy != 0 || pxz != 0 || pyz != 0) {
        float od = 0.f;
        bool first = true;
        if constexpr (pxy != 0) { od = r.M[5]; first = false; }
        if constexpr (pxz != 0) { od = first ? r.M[6] : (pxz * s_off > 0 ? od + r.M[6] : od - r.M[6]); first = false; }
        if constexpr (pyz != 0) { od = first ? r.M[8] : (pyz * s_off > 0 ? od + r.M[8] : od - r.M[8]); }
        e = s_off > 0 ? e + od : e - od;
    }
    // odd part, canonical sign
    constexpr int s_odd = x != 0 ? x : (y != 0 ? y : (z != 0 ? z : 1));
    float h = e;
    if constexpr (x != 0 || y != 0 || z != 0) {
        float o = 0.f;
        bool first = true;
        if constexpr (x != 0) { o = r.M[1]; first = false; }
        if constexpr (y != 0) { o = first ? r.M[2] : (y * s_odd > 0 ? o + r.M[2] : o - r.M[2]); first = false; }
        if constexpr (z != 0) { o = first ? r.M[3] : (z * s_odd > 0 ? o + r.M[3] : o - r.M[3]); }
        h = s_odd > 0 ? e + o : e - o;
    }
    constexpr int W = wint<Q>(q);
    if constexpr (W == 1)
        return h;
    else
        return float(W) * h;
}

// Raw second-order sums of H, accumulated over opposite pairs (q, q+1):
// P = H_q + H_q+1 feeds the even sums, Mi = H_q - H_q+1 the odd ones.
struct Sums {
    float s0, sx, sy, sz, sxx, sxy, sxz, syy, syz, szz;
};

template <int Q>
__device__ __forceinline__ Sums pop_sums(const float (&g)[27]) {
    Sums s{g[0], 0, 0, 0, 0, 0, 0, 0, 0, 0};
    unroll<0, (Q - 1) / 2>([&](auto pc) {
        constexpr int q = 1 + 2 * decltype(pc)::value;  // q and q+1 are opposite
        constexpr int x = cxq(q), y = cyq(q), z = czq(q);
        static_assert(cxq(q + 1) == -x && cyq(q + 1) == -y && czq(q + 1) == -z, "pair order");
        const float P = g[q] + g[q + 1];
        const float Mi = g[q] - g[q + 1];
        s.s0 += P;
        if constexpr (x != 0) s.sx = x > 0 ? s.sx + Mi : s.sx - Mi;
        if constexpr (y != 0) s.sy = y > 0 ? s.sy + Mi : s.sy - Mi;
        if constexpr (z != 0) s.sz = z > 0 ? s.sz + Mi : s.sz - Mi;
        if constexpr (x != 0) s.sxx += P;
        if constexpr (y != 0) s.syy += P;
        if constexpr (z != 0) s.szz += P;
        if constexpr (x * y != 0) s.sxy = x * y > 0 ? s.sxy + P : s.sxy - P;
        if constexpr (x * z != 0) s.sxz = x * z > 0 ? s.sxz + P : s.sxz - P;
        if constexpr (y * z != 0) s.syz = y * z > 0 ? s.syz + P : s.syz - P;
    });
    return s;
}

// Interior moments (velocitySet.cuh:398-417) followed by scale (velocitySet.cuh:130-142):
// rho' = S0/3D, R = 3D rho = 3D + S0, 3u = 3 Sx/R, mxx = (4.5 Sxx - 1.5 S0)/R, mxy = 9 Sxy/R.
// x / 3D without a systematic bias: 1/3D split into a float head and tail
// (together exact to ~2^-48), so the product carries at most ~1 ulp of
// round-off that does not repeat identically step after step.
template <int Q>
__device__ __forceinline__ float div_d3(float x) {
    constexpr double inv = Q == 19 ? 1.0 / 108.0 : 1.0 / 648.0;
    constexpr float hi = float(inv);
    constexpr float lo = float(inv - double(hi));
    return fmaf(x, hi, x * lo);
}
// 1/r from the hardware reciprocal plus one Newton step (~0.5 ulp).
__device__ __forceinline__ float recip(float r) {
    float y = __builtin_amdgcn_rcpf(r);
    const float e = fmaf(-r, y, 1.0f);
    return fmaf(y, e, y);
}

template <int Q>
__device__ __forceinline__ void interior_moments(const Sums &s, float (&m)[10]) {
    const float inv = recip(d3<Q>() + s.s0);
    m[0] = div_d3<Q>(s.s0);
    m[1] = 3.0f * s.sx * inv;
    m[2] = 3.0f * s.sy * inv;
    m[3] = 3.0f * s.sz * inv;
    m[4] = fmaf(4.5f, s.sxx, -1.5f * s.s0) * inv;
    m[5] = 9.0f * s.sxy * inv;
    m[6] = 9.0f * s.sxz * inv;
    m[7] = fmaf(4.5f, s.syy, -1.5f * s.s0) * inv;
    m[8] = 9.0f * s.syz * inv;
    m[9] = fmaf(4.5f, s.szz, -1.5f * s.s0) * inv;
}

// Boundary moments: boundaryConditions.cuh:99-554 rewritten in deviation form.
// With rho_I = W + rI and (m_ab,I rho_I) = A + a_ab, where W and A are the known
// weight sums of the node type (exact rationals) and rI, a_ab the known sums of
// g = f - w, each closure rho = (alpha rho_I + beta a)/gamma becomes
// rho - 1 = (alpha rI + beta a + c0)/gamma with c0 = alpha W + beta A - gamma,
// which is identically zero for every static wall, edge and corner (their
// closures keep the rest state), so it is dropped there and kept in closed form
// for the lid.  `bits` is the normalVector bitmask (W=1, E=2, S=4, N=8, B=16, F=32).
template <int Q>
__device__ __forceinline__ void boundary_moments(const Sums &s, int bits, float omega, float U,
                                                 float (&m)[10]) {
    // divisions as reciprocal products (recip: hardware reciprocal + one Newton
    // step, ~0.5 ulp; div_d3: the two-constant product): these lanes run the
    // closure beside the interior lanes of their wave, so its cost is paid by
    // every wave with a boundary node in it
    const float rI = div_d3<Q>(s.s0);
    const float axy = div_d3<Q>(s.sxy), axz = div_d3<Q>(s.sxz), ayz = div_d3<Q>(s.syz);
    const float w = omega, tw = 1.0f - omega;
    const bool north = (bits & 8) != 0;
    const int nb = __popc(bits & 0x3F);
    const float ux = north ? U * (nb == 1 ? 1.f : (nb == 2 ? 0.5f : 1.f / 3.f)) : 0.f;  // U<X> (boundaryConditions.cuh:543-554)
    float rp = 0.f, mxy = 0.f, mxz = 0.f, myz = 0.f;
    // static edge: sign +1 is rho = 36(rho_I - a + a w)/(24 + w), mab = (36a - rho)/(9 rho);
    // sign -1 is rho = -36(-rho_I - a + a w)/(24 + w), mab = (36a + rho)/(9 rho)
    auto edge = [&](float sign, float a, float &mab) {
        rp = 36.f * (rI - sign * tw * a) * recip(24.f + w);
        mab = (36.f * a - sign * rp) * recip(9.f * (1.f + rp));
    };
    switch (bits & 0x3F) {
    case 21: case 37: case 22: case 38:  // SWB SWF SEB SEF: rho = 12 rho_I/7 (216/125 for D3Q27)
        rp = rI * ((Q == 19) ? 12.f / 7.f : 216.f / 125.f);
        break;
    case 5: edge(+1.f, axy, mxy); break;   // SW
    case 6: edge(-1.f, axy, mxy); break;   // SE
    case 17: edge(+1.f, axz, mxz); break;  // WB
    case 33: edge(-1.f, axz, mxz); break;  // WF
    case 18: edge(-1.f, axz, mxz); break;  // EB
    case 34: edge(+1.f, axz, mxz); break;  // EF
    case 20: edge(+1.f, ayz, myz); break;  // SB
    case 36: edge(-1.f, ayz, myz); break;  // SF
    case 1: case 2: {  // W, E: rho = 6 rho_I/5, mab = 2 a/rho
        rp = 1.2f * rI;
        const float ir = recip(1.f + rp);
        mxy = 2.f * axy * ir;
        mxz = 2.f * axz * ir;
    } break;
    case 4: {  // S
        rp = 1.2f * rI;
        const float ir = recip(1.f + rp);
        mxy = 2.f * axy * ir;
        myz = 2.f * ayz * ir;
    } break;
    case 16: case 32: {  // B, F
        rp = 1.2f * rI;
        const float ir = recip(1.f + rp);
        mxz = 2.f * axz * ir;
        myz = 2.f * ayz * ir;
    } break;
    case 8: {  // N (moving lid): mxy = (6a - U rho)/(3 rho), myz = 2a/rho
        rp = 1.2f * rI;
        const float rho = 1.f + rp;
        const float ir = recip(rho);
        mxy = (6.f * axy - U * rho) * (ir * (1.f / 3.f));
        myz = 2.f * ayz * ir;
    } break;
    case 25: case 41: {  // NWB NWF: rho = 24 rho_I/(14 + 8U - 9U^2); W = 7/12 (125/216 for D3Q27)
        const float c0 = (Q == 19) ? 0.f : -1.f / 9.f;
        rp = (24.f * rI + c0 - 8.f * U + 9.f * U * U) * recip(14.f + 8.f * U - 9.f * U * U);
    } break;
    case 26: case 42: {  // NEB NEF: rho = 24 rho_I/(14 - 8U - 9U^2)
        const float c0 = (Q == 19) ? 0.f : -1.f / 9.f;
        rp = (24.f * rI + c0 + 8.f * U + 9.f * U * U) * recip(14.f - 8.f * U - 9.f * U * U);
    } break;
    case 24: {  // NB: rho = 72(-rho_I - a + a w)/(-48 - 2w + 3U^2 w)
        const float g = -48.f - 2.f * w + 3.f * U * U * w;
        rp = (-72.f * rI - 72.f * tw * ayz - 3.f * U * U * w) * recip(g);
        const float rho = 1.f + rp;
        myz = (72.f * ayz + 2.f * rp - 3.f * U * U * rho) * recip(18.f * rho);
    } break;
    case 40: {  // NF: rho = -72(rho_I - a + a w)/(-48 - 2w + 3U^2 w)
        const float g = -48.f - 2.f * w + 3.f * U * U * w;
        rp = (-72.f * rI + 72.f * tw * ayz - 3.f * U * U * w) * recip(g);
        const float rho = 1.f + rp;
        myz = (72.f * ayz - 2.f * rp + 3.f * U * U * rho) * recip(18.f * rho);
    } break;
    case 10: {  // NE: rho = 36(rho_I - a + a w)/g
        const float g = 24.f - 18.f * U - 18.f * U * U + w + 3.f * U * w + 3.f * U * U * w;
        const float c0 = 18.f * U + 18.f * U * U - 3.f * U * w - 3.f * U * U * w;  // 24 + w - g
        rp = (36.f * rI - 36.f * tw * axy + c0) * recip(g);
        const float rho = 1.f + rp;
        mxy = (36.f * axy - rp - 3.f * U * rho - 3.f * U * U * rho) * recip(9.f * rho);
    } break;
    case 9: {  // NW: rho = -36(-rho_I - a + a w)/g
        const float g = 24.f + 18.f * U - 18.f * U * U + w - 3.f * U * w + 3.f * U * U * w;
        const float c0 = -18.f * U + 18.f * U * U + 3.f * U * w - 3.f * U * U * w;  // 24 + w - g
        rp = (36.f * rI + 36.f * tw * axy + c0) * recip(g);
        const float rho = 1.f + rp;
        mxy = (36.f * axy + rp - 3.f * U * rho + 3.f * U * U * rho) * recip(9.f * rho);
    } break;
    default: break;
    }
    // unscaled -> scaled (velocitySet.cuh:130-142); m_xx = u_x^2, m_yy = m_zz = 0
    m[0] = rp;
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
