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

// Floating-point contraction is off for everything in this file: a multiply
// and an add fuse into an FMA only where the source says fmaf, so every kernel
// that inlines these functions (the step of any tiling, the batched step, the
// face rebuild) rounds exactly alike and a restart or another tiling
// reproduces the same bits.
#pragma float_control(push)
#pragma clang fp contract(off)
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

// Up-going populations that also move in-plane (c_z = +1, q != 5), and the
// index of one among them (ascending q).
__host__ __device__ constexpr bool up_moving(int q) { return czq(q) == 1 && (cxq(q) != 0 || cyq(q) != 0); }
template <int Q>
__host__ __device__ constexpr int u_index(int q) {
    int n = 0;
    for (int p = 1; p < q; p++)
        if (up_moving(p)) n++;
    return n;
}
// LDS slot of a population that moves in-plane (q not in {0,5,6}): the others
// first in ascending q, the NU up-going ones last in u_index order, so that
// their second bank (slots NSLOT + u_index) is the first bank shifted by NU
// slots for every one of them.
template <int Q>
__host__ __device__ constexpr int lds_slot(int q) {
    if (up_moving(q)) return VS<Q>::NSLOT - VS<Q>::NU + u_index<Q>(q);
    int n = 0;
    for (int p = 1; p < q; p++)
        if (p != 5 && p != 6 && !up_moving(p)) n++;
    return n;
}

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
__host__ __device__ constexpr int opposite(int q) { return q == 0 ? 0 : ((q & 1) ? q + 1 : q - 1); }
// Sign of the first nonzero component of c_q: H_q = W (even + odd_sign * odd).
template <int q>
__host__ __device__ constexpr int odd_sign() {
    return cxq(q) != 0 ? cxq(q) : (cyq(q) != 0 ? cyq(q) : (czq(q) != 0 ? czq(q) : 1));
}
// Even part: a + the diagonal terms, then the off-diagonal sum with its
// canonical sign (first nonzero product positive); equal for q and opposite(q).
template <int Q, int q>
__device__ __forceinline__ float rec_even(const Rec &r) {
    constexpr int x = cxq(q), y = cyq(q), z = czq(q);
    float e = r.a;
    if constexpr (x != 0) e += r.M[4];
    if constexpr (y != 0) e += r.M[7];
    if constexpr (z != 0) e += r.M[9];
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
    return e;
}
// Odd part with the canonical sign (q != 0); equal for q and opposite(q).
template <int q>
__device__ __forceinline__ float rec_odd(const Rec &r) {
    constexpr int x = cxq(q), y = cyq(q), z = czq(q);
    constexpr int s_odd = odd_sign<q>();
    float o = 0.f;
    bool first = true;
    if constexpr (x != 0) { o = r.M[1]; first = false; }
    if constexpr (y != 0) { o = first ? r.M[2] : (y * s_odd > 0 ? o + r.M[2] : o - r.M[2]); first = false; }
    if constexpr (z != 0) { o = first ? r.M[3] : (z * s_odd > 0 ? o + r.M[3] : o - r.M[3]); }
    return o;
}
// D3Q27 populations travel (LDS, face buffers) as h_q = H_q / Wint_q, the
// weight (64, 16, 4 or 1: a power of two) applied once inside the moment sums
// (pop_sums) as the multiplier of an fma that is there anyway; the sums round
// exactly as they would from H_q.  The instruction-bound D3Q27 FP16 step gains
// 2-3% (256^3 170.1 -> 165.8 us); for D3Q19, whose rest weight 12 is not a
// power of two, the same change measured slower (FP32 246.9 -> 249.3 us).
template <int Q, int q>
__device__ __forceinline__ float rec_dev(const Rec &r) {
    const float e = rec_even<Q, q>(r);
    float h = e;
    if constexpr (q != 0) {
        const float o = rec_odd<q>(r);
        h = odd_sign<q>() > 0 ? e + o : e - o;
    }
    constexpr int W = wint<Q>(q);
    if constexpr (W == 1 || Q == 27)
        return h;
    else
        return float(W) * h;
}
// H_q for sgn = +1 and H_opposite(q) for sgn = -1 (a per-lane choice between the
// two populations of a pair), with the same bits as rec_dev: e +/- o rounded once.
template <int Q, int q>
__device__ __forceinline__ float rec_signed(const Rec &r, float sgn) {
    static_assert(q != 0, "the rest population has no pair");
    const float e = rec_even<Q, q>(r);
    const float o = rec_odd<q>(r);
    const float h = odd_sign<q>() > 0 ? fmaf(sgn, o, e) : fmaf(-sgn, o, e);
    constexpr int W = wint<Q>(q);
    if constexpr (W == 1 || Q == 27)
        return h;
    else
        return float(W) * h;
}

// Raw second-order sums of H, accumulated over opposite pairs (q, q+1):
// P = H_q + H_q+1 feeds the even sums, Mi = H_q - H_q+1 the odd ones.
struct Sums {
    float s0, sx, sy, sz, sxx, sxy, sxz, syy, syz, szz;
};

// Raw sums of the pulled populations, formed from the opposite pairs
// (q, q+1): P = H_q + H_q+1 feeds the even sums, M = H_q - H_q+1 the odd ones.
// The pairs are grouped so that every sub-sum serves several moments: the two
// edge pairs of a coordinate plane as their sum (diagonal moments, density)
// and difference (off-diagonal moment), and for D3Q27 the four corner pairs
// through four half-sums/differences from which the corners' share of every
// moment follows with one more operation -- 48 adds for D3Q19 and 81 for
// D3Q27 against 54 and 102 when the pairs are accumulated one by one.  D3Q27's
// g holds h = H / Wint (see rec_dev): its weights enter here as fma multipliers.
template <int q, int x, int y, int z>
__host__ __device__ constexpr bool is_c() { return cxq(q) == x && cyq(q) == y && czq(q) == z; }
template <int Q>
__device__ __forceinline__ Sums pop_sums(const float (&g)[27]) {
    static_assert(is_c<1, 1, 0, 0>() && is_c<3, 0, 1, 0>() && is_c<5, 0, 0, 1>() && is_c<7, 1, 1, 0>() &&
                      is_c<9, 1, 0, 1>() && is_c<11, 0, 1, 1>() && is_c<13, 1, -1, 0>() && is_c<15, 1, 0, -1>() &&
                      is_c<17, 0, 1, -1>(),
                  "axis and edge pair order");
    auto P = [&](int q) { return g[q] + g[q + 1]; };
    auto M = [&](int q) { return g[q] - g[q + 1]; };
    // axis pairs
    const float Ax = P(1), Ay = P(3), Az = P(5), Mx = M(1), My = M(3), Mz = M(5);
    // edge pairs: xy (q 7: (1,1,0), q 13: (1,-1,0)), xz (9, 15), yz (11, 17)
    const float Pxy = P(7), Qxy = P(13), Pxz = P(9), Qxz = P(15), Pyz = P(11), Qyz = P(17);
    const float Nxy = M(7), Oxy = M(13), Nxz = M(9), Oxz = M(15), Nyz = M(11), Oyz = M(17);
    const float Exy = Pxy + Qxy, Exz = Pxz + Qxz, Eyz = Pyz + Qyz;  // -> diagonal sums, density
    const float Dxy = Pxy - Qxy, Dxz = Pxz - Qxz, Dyz = Pyz - Qyz;  // -> off-diagonal sums
    // odd edge parts: x from both xy and both xz pairs, y from xy (sign of c_y) and yz, z from xz and yz
    const float Sxy = Nxy + Oxy, Txy = Nxy - Oxy, Sxz = Nxz + Oxz, Txz = Nxz - Oxz, Syz = Nyz + Oyz,
                Tyz = Nyz - Oyz;
    Sums s;
    if constexpr (Q == 19) {
        s.sxx = (Ax + Exy) + Exz;
        s.syy = (Ay + Exy) + Eyz;
        s.szz = (Az + Exz) + Eyz;
        s.s0 = g[0] + ((((((Ax + Ay) + Az) + Exy) + Exz) + Eyz));
        s.sxy = Dxy;
        s.sxz = Dxz;
        s.syz = Dyz;
        s.sx = (Mx + Sxy) + Sxz;
        s.sy = (My + Txy) + Syz;
        s.sz = (Mz + Txz) + Tyz;
    } else {
        // corner pairs 19 (1,1,1), 21 (1,1,-1), 23 (1,-1,1), 25 (-1,1,1)
        static_assert(is_c<19, 1, 1, 1>() && is_c<21, 1, 1, -1>() && is_c<23, 1, -1, 1>() && is_c<25, -1, 1, 1>(),
                      "corner pair order");
        static_assert(wint<27>(0) == 64 && wint<27>(1) == 16 && wint<27>(7) == 4 && wint<27>(19) == 1, "D3Q27 weights");
        const float P1 = P(19), P2 = P(21), P3 = P(23), P4 = P(25);
        const float M1 = M(19), M2 = M(21), M3 = M(23), M4 = M(25);
        const float a = P1 + P2, b = P3 + P4, c = P1 - P2, d = P3 - P4;
        const float C = a + b;                          // every corner: diagonal sums, density
        const float Cxy = a - b, Cxz = c + d, Cyz = c - d;  // signs of c_x c_y, c_x c_z, c_y c_z
        const float e = M1 + M2, f = M3 - M4, h = M1 - M2, k = M3 + M4;
        // weights relative to the next class: axis 16 = 4 x edge 4 = 4 x corner 1;
        // (4 Ax + E) = (16 Ax + 4 E) / 4 exactly, so each fma rounds as the sum
        // of the weighted terms would
        const float A3 = (Ax + Ay) + Az;
        s.sxx = fmaf(4.f, fmaf(4.f, Ax, Exy) + Exz, C);
        s.syy = fmaf(4.f, fmaf(4.f, Ay, Exy) + Eyz, C);
        s.szz = fmaf(4.f, fmaf(4.f, Az, Exz) + Eyz, C);
        s.s0 = fmaf(64.f, g[0], fmaf(4.f, (fmaf(4.f, A3, Exy) + Exz) + Eyz, C));
        s.sxy = fmaf(4.f, Dxy, Cxy);
        s.sxz = fmaf(4.f, Dxz, Cxz);
        s.syz = fmaf(4.f, Dyz, Cyz);
        s.sx = fmaf(4.f, fmaf(4.f, Mx, Sxy) + Sxz, e + f);  // corners' c_x: +, +, +, -
        s.sy = fmaf(4.f, fmaf(4.f, My, Txy) + Syz, e - f);  // c_y: +, +, -, +
        s.sz = fmaf(4.f, fmaf(4.f, Mz, Txz) + Tyz, h + k);  // c_z: +, -, +, +
    }
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
    const float inv3 = 3.0f * inv, inv9 = 9.0f * inv;
    m[0] = div_d3<Q>(s.s0);
    m[1] = s.sx * inv3;
    m[2] = s.sy * inv3;
    m[3] = s.sz * inv3;
    m[4] = fmaf(4.5f, s.sxx, -1.5f * s.s0) * inv;
    m[5] = s.sxy * inv9;
    m[6] = s.sxz * inv9;
    m[7] = fmaf(4.5f, s.syy, -1.5f * s.s0) * inv;
    m[8] = s.syz * inv9;
    m[9] = fmaf(4.5f, s.szz, -1.5f * s.s0) * inv;
}

// Boundary moments: boundaryConditions.cuh:99-554 rewritten in deviation form.
// With rho_I = W + rI and (m_ab,I rho_I) = A + a_ab, where W and A are the known
// weight sums of the node type (exact rationals) and rI, a_ab the known sums of
// g = f - w, each closure rho = (alpha rho_I + beta a)/gamma becomes
// rho - 1 = (alpha rI + beta a + c0)/gamma with c0 = alpha W + beta A - gamma,
// which is identically zero for every static wall, edge and corner (their
// closures keep the rest state), so it is dropped there and kept in closed form
// for the lid.  `code` is the normalVector bitmask without its boundary bit
// (W=1, E=2, S=4, N=8, B=16, F=32).
//
// All 26 closures share one form, so they are evaluated from a per-case
// coefficient row instead of a switch (a wave holding lanes of several cases
// would otherwise run each case's branch in turn, next to its interior lanes):
//   rho - 1 = (A rI + Bxy axy + Bxz axz + Byz ayz + C) / D
//   m_k     = (E_k a_k + F_k (rho - 1) + G_k rho) / (H_k rho),  k = xy, xz, yz
// with the stored velocity u_x = U / (number of boundaries) on lid nodes
// (boundaryConditions.cuh:543-554) and m_xx = u_x^2, m_yy = m_zz = 0.
struct Closure {
    float A, Bxy, Bxz, Byz, C, invD, ux, pad;
    float E[3], F[3], G[3], invH[3];
};

// Row of a node code in a 27-row closure table: one ternary digit per axis
// (0: inside, 1: W/S/B wall, 2: E/N/F wall).
__host__ __device__ __forceinline__ int closure_row(int code) {
    const int ix = (code & 1) ? 1 : ((code & 2) ? 2 : 0);
    const int iy = (code & 4) ? 1 : ((code & 8) ? 2 : 0);
    const int iz = (code & 16) ? 1 : ((code & 32) ? 2 : 0);
    return ix + 3 * iy + 9 * iz;
}
__host__ __device__ __forceinline__ int closure_code(int row) {
    const int ix = row % 3, iy = (row / 3) % 3, iz = row / 9;
    return (ix == 1 ? 1 : ix == 2 ? 2 : 0) | (iy == 1 ? 4 : iy == 2 ? 8 : 0) | (iz == 1 ? 16 : iz == 2 ? 32 : 0);
}

template <int Q>
__device__ inline Closure closure_coeffs(int code, float U, float w) {
    Closure k{};
    k.invH[0] = k.invH[1] = k.invH[2] = 1.f;
    k.invD = 1.f;
    const float tw = 1.0f - w, U2 = U * U;
    const int nb = __popc(code);
    k.ux = (code & 8) ? U / float(nb) : 0.f;
    // static edge: rho = 36(rho_I - s tw a)/(24 + w), m = (36 a - s rho')/(9 rho)
    auto edge = [&](float sign, int a) {
        k.A = 36.f;
        (a == 0 ? k.Bxy : a == 1 ? k.Bxz : k.Byz) = -36.f * sign * tw;
        k.invD = 1.f / (24.f + w);
        k.E[a] = 36.f;
        k.F[a] = -sign;
        k.invH[a] = 1.f / 9.f;
    };
    switch (code) {
    case 21: case 37: case 22: case 38:  // SWB SWF SEB SEF: rho = 12 rho_I/7 (216/125 for D3Q27)
        k.A = (Q == 19) ? 12.f / 7.f : 216.f / 125.f;
        break;
    case 5: edge(+1.f, 0); break;   // SW
    case 6: edge(-1.f, 0); break;   // SE
    case 17: edge(+1.f, 1); break;  // WB
    case 33: edge(-1.f, 1); break;  // WF
    case 18: edge(-1.f, 1); break;  // EB
    case 34: edge(+1.f, 1); break;  // EF
    case 20: edge(+1.f, 2); break;  // SB
    case 36: edge(-1.f, 2); break;  // SF
    case 1: case 2:  // W, E: rho = 6 rho_I/5, mab = 2 a/rho
        k.A = 1.2f;
        k.E[0] = k.E[1] = 2.f;
        break;
    case 4:  // S
        k.A = 1.2f;
        k.E[0] = k.E[2] = 2.f;
        break;
    case 16: case 32:  // B, F
        k.A = 1.2f;
        k.E[1] = k.E[2] = 2.f;
        break;
    case 8:  // N (moving lid): mxy = (6a - U rho)/(3 rho), myz = 2a/rho
        k.A = 1.2f;
        k.E[0] = 6.f;
        k.G[0] = -U;
        k.invH[0] = 1.f / 3.f;
        k.E[2] = 2.f;
        break;
    case 25: case 41: {  // NWB NWF: rho = 24 rho_I/(14 + 8U - 9U^2); W = 7/12 (125/216 for D3Q27)
        k.A = 24.f;
        k.C = ((Q == 19) ? 0.f : -1.f / 9.f) - 8.f * U + 9.f * U2;
        k.invD = 1.f / (14.f + 8.f * U - 9.f * U2);
    } break;
    case 26: case 42: {  // NEB NEF: rho = 24 rho_I/(14 - 8U - 9U^2)
        k.A = 24.f;
        k.C = ((Q == 19) ? 0.f : -1.f / 9.f) + 8.f * U + 9.f * U2;
        k.invD = 1.f / (14.f - 8.f * U - 9.f * U2);
    } break;
    case 24: case 40: {  // NB, NF: rho = -/+72(rho_I -/+ a + a w)/(-48 - 2w + 3U^2 w)
        const float s = code == 24 ? 1.f : -1.f;
        k.A = -72.f;
        k.Byz = -72.f * s * tw;
        k.C = -3.f * U2 * w;
        k.invD = 1.f / (-48.f - 2.f * w + 3.f * U2 * w);
        k.E[2] = 72.f;
        k.F[2] = 2.f * s;
        k.G[2] = -3.f * U2 * s;
        k.invH[2] = 1.f / 18.f;
    } break;
    case 10: case 9: {  // NE, NW: rho = 36(rho_I -/+ a +/- a w)/g
        const float s = code == 10 ? 1.f : -1.f;  // NE: +1
        const float g = 24.f - 18.f * s * U - 18.f * U2 + w + 3.f * s * U * w + 3.f * U2 * w;
        k.A = 36.f;
        k.Bxy = -36.f * s * tw;
        k.C = 18.f * s * U + 18.f * U2 - 3.f * s * U * w - 3.f * U2 * w;  // 24 + w - g
        k.invD = 1.f / g;
        k.E[0] = 36.f;
        k.F[0] = -s;
        k.G[0] = -3.f * U - 3.f * s * U2;
        k.invH[0] = 1.f / 9.f;
    } break;
    default: break;
    }
    return k;
}

// Moments of a boundary node from its known sums and its closure row
// (reciprocals: hardware reciprocal + one Newton step; 1/(3D) as a two-term
// product), scaled as stored (velocitySet.cuh:130-142).
template <int Q>
__device__ __forceinline__ void boundary_moments(const Sums &s, const Closure &k, float (&m)[10]) {
    const float rI = div_d3<Q>(s.s0);
    const float a[3] = {div_d3<Q>(s.sxy), div_d3<Q>(s.sxz), div_d3<Q>(s.syz)};
    const float rp = fmaf(k.A, rI, fmaf(k.Bxy, a[0], fmaf(k.Bxz, a[1], fmaf(k.Byz, a[2], k.C)))) * k.invD;
    const float rho = 1.f + rp;
    const float ir = recip(rho);
    float mab[3];
#pragma unroll
    for (int j = 0; j < 3; j++) mab[j] = fmaf(k.E[j], a[j], fmaf(k.F[j], rp, k.G[j] * rho)) * (k.invH[j] * ir);
    m[0] = rp;
    m[1] = 3.f * k.ux;
    m[2] = 0.f;
    m[3] = 0.f;
    m[4] = 4.5f * k.ux * k.ux;
    m[5] = 9.f * mab[0];
    m[6] = 9.f * mab[1];
    m[7] = 0.f;
    m[8] = 9.f * mab[2];
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
#pragma float_control(pop)
