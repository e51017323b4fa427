// step_kernel.hpp -- the fused moment-representation LBM step for gfx950.
//
// One workgroup owns a TX x TY column footprint of the domain and marches
// through `lz` consecutive z-planes of it (a "chunk").  Per plane it
//   1. holds the old moments of planes z and z+1 in registers (z+2 is in flight),
//   2. rebuilds from them the populations that stream in-plane (cz = 0), down
//      (cz = -1, from plane z+1) and -- kept from the previous plane -- up
//      (cz = +1, from plane z-1), and stages them in LDS,
//   3. fills a one-node LDS ring around the footprint from the neighbour
//      columns' face buffers (written in the previous step),
//   4. gathers every population from its source node x - c_q,
//   5. forms moments (interior) or applies the wall/lid conditions (boundary),
//      collides, writes the 10 moments back in place,
//   6. writes the outgoing populations of its footprint edges to the other
//      face-buffer set (x/y faces every plane, z faces at the chunk ends).
// The face buffers therefore cost 4*QF*4 bytes*(1/TX + 1/TY + 1/lz) per node
// instead of the 8^3-tile scheme's 4*QF*4*(3/8) (reference halo.cuh:171-494).
//
// Semantics follow the reference step (momentBasedD3Q19.cuh:117-184): pull
// streaming f_q(x) = f*_q(x - c_q) of post-collision populations rebuilt from
// the stored moments, boundary moments from the known populations, scale,
// collide, store.
#pragma once

#include <hip/hip_fp16.h>

#include "lattice.hpp"

// Floating-point contraction is off for everything in this file: a multiply
// and an add fuse into an FMA only where the source says fmaf, so every kernel
// that inlines these functions (the step of any tiling, the batched step, the
// face rebuild) rounds exactly alike and a restart or another tiling
// reproduces the same bits.
#pragma float_control(push)
#pragma clang fp contract(off)
namespace lbm {

// Barrier that orders LDS traffic only: global loads issued for the next plane
// stay in flight across it (a plain __syncthreads would drain them).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Step counter read by the fused time means and the FP16 stochastic rounding:
// t of the step about to run and invNewCount = 1/(t+1) (kineticEnergy.cuh:249).
// Two of them alternate with the face-set parity: a step reads its own and
// one thread of the step writes the next step's (t+1), which no other
// workgroup of the same launch reads -- no separate kernel per step.
struct StepClock {
    int64_t t;
    double inv;
};

struct StepParams {
    int nx, ny, nz, lz, nbx, nby, nc;
    float omega, t_omega, omega_h, u_lid, tau;
    // Stored moments, column-tiled internal layout, in place, in 4-byte node
    // words (one FP32 moment, or the FP16 codes of moments 2j and 2j+1, moment
    // 2j+1 at mom[2j] + 2 bytes).  Node i of the internal order (plane block
    // b = i / (tx ty), r = i mod (tx ty)) of moment k is in word b * mplane + r
    // from mom[k]: mplane = tx ty for separate word arrays, (number of word
    // arrays) tx ty when they are interleaved per plane block (every plane
    // block of a workgroup then is one contiguous run of all ten moments).
    // mshift = log2(tx ty).
    void *mom[10];
    int mplane, mshift;
    const float *fin;      // face buffers read (previous step): X | Y | Z
    float *fout;           // face buffers written (this step)
    int yoff, zoff;        // offsets of the Y and Z face blocks inside one face set
    float enc[10], dec[10];  // codec_enc/codec_dec (FP16) or 1 (FP32), for the auxiliary kernels
    // fused FP64 time means of the function objects (fo_mask bits 2: kMean,
    // 8: SMean, 16: momentsMean)
    unsigned fo_mask;
    // The enabled means' components (kMean; SMean xx..zz; momentsMean of the ten
    // stored moments) are interleaved per plane block in one allocation:
    // component c of node r of plane block b at b * fplane + c * fcomp + r, with
    // fcomp = tx ty and fplane = (number of enabled components) tx ty.  kmean,
    // smean, mmean point at the first component of each mean (null when off).
    double *kmean;
    double *smean;
    double *mmean;
    int fplane, fcomp;
    const StepClock *clock; // absolute step index of this step and 1/(t+1) (device)
    StepClock *clock_next;  // the next step's clock, or null when nothing reads it
    float s_diag, s_off;    // strain-rate coefficients 1/(3 tau), 1/(6 tau)
    int64_t n;             // nodes
    // probe outputs
    uint32_t *probe_src;
    uint8_t *probe_type;
};

// FP16 storage codec: stored16 = m * codec_enc(k), m = stored16 * codec_dec(k);
// (rho-1)/cs2, u/cs, (Pi_ab/rho - cs2 d_ab)/cs2 with the stored 4.5/9 factors
// folded in.  Compile-time constants, so the step holds no codec registers.
__host__ __device__ constexpr double codec_enc_d(int k) {
    return k == 0 ? 3.0 : k <= 3 ? 0.57735026918962576451 : (k == 4 || k == 7 || k == 9) ? 2.0 / 3.0 : 1.0 / 3.0;
}
__host__ __device__ constexpr float codec_enc(int k) { return (float)codec_enc_d(k); }
__host__ __device__ constexpr float codec_dec(int k) { return (float)(1.0 / codec_enc_d(k)); }

template <typename ST>
struct Store;
// Element i of a moment array is addressed as (array base) + (32-bit byte
// offset): the base is a kernel argument in scalar registers, so a load or
// store maps to the scalar-base + vector-offset form of the global memory
// instructions, and one offset per node serves all ten arrays (no 64-bit
// address arithmetic per access).  The host keeps every moment array under
// 4 GiB (at most 2^30 nodes).
__device__ __forceinline__ const void *byte_at(const void *p, unsigned off) {
    return static_cast<const char *>(p) + off;
}
__device__ __forceinline__ void *byte_at(void *p, unsigned off) { return static_cast<char *>(p) + off; }
// Moment arrays are streamed once per step with the nontemporal cache policy,
// which leaves the caches to the face buffers that the next step reads back
// (face stores with the same policy made the FP32 D3Q19 step 8% slower).
template <>
struct Store<float> {
    static constexpr unsigned kBytes = 4;
    static constexpr int kWords = 10;  // 4-byte words of one node's moments
    __device__ static __forceinline__ float load(const void *p, unsigned off, float) {
        return __builtin_nontemporal_load(static_cast<const float *>(byte_at(p, off)));
    }
    __device__ static __forceinline__ float load_raw(const void *p, unsigned off) { return load(p, off, 1.f); }
    __device__ static __forceinline__ float raw_word(float v) { return v; }
    __device__ static __forceinline__ void store(void *p, unsigned off, float v) {
        __builtin_nontemporal_store(v, static_cast<float *>(byte_at(p, off)));
    }
    __device__ static __forceinline__ void store_word(void *p, unsigned off, uint32_t w) {
        __builtin_nontemporal_store(w, static_cast<uint32_t *>(byte_at(p, off)));
    }
};
// FP16: the codes of moments 2j (low half) and 2j + 1 (high half) of a node
// share one 32-bit node word (mom[2j + 1] = mom[2j] + 2 bytes), so a plane of
// ten codes is five 4-byte loads and stores per node; kBytes is the node stride.
template <>
struct Store<_Float16> {
    static constexpr unsigned kBytes = 4;
    static constexpr int kWords = 5;  // 4-byte node words of one node's moments
    __device__ static __forceinline__ float load(const void *p, unsigned off, float dec) {
        return float(__builtin_nontemporal_load(static_cast<const _Float16 *>(byte_at(p, off)))) * dec;
    }
    // The stored code as a raw word (the 16-bit code in the low half of a
    // 32-bit word, as the load leaves it): rec_setup_st decodes it with
    // mixed-precision FMAs that take the FP16 operand directly.
    __device__ static __forceinline__ float load_raw(const void *p, unsigned off) {
        return raw_word(__builtin_nontemporal_load(static_cast<const _Float16 *>(byte_at(p, off))));
    }
    __device__ static __forceinline__ float raw_word(_Float16 v) {
        return __uint_as_float((uint32_t)__builtin_bit_cast(uint16_t, v));
    }
    __device__ static __forceinline__ float from_raw(float w) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)__float_as_uint(w));
    }
    __device__ static __forceinline__ float from_raw_hi(float w) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(__float_as_uint(w) >> 16));
    }
    __device__ static __forceinline__ void store(void *p, unsigned off, _Float16 v) {
        __builtin_nontemporal_store(v, static_cast<_Float16 *>(byte_at(p, off)));
    }
    // the node word of moments 2j, 2j + 1 (p = mom[2j]), as float bits
    __device__ static __forceinline__ float load_word(const void *p, unsigned off) {
        return __uint_as_float(__builtin_nontemporal_load(static_cast<const uint32_t *>(byte_at(p, off))));
    }
    __device__ static __forceinline__ void store_word(void *p, unsigned off, uint32_t w) {
        __builtin_nontemporal_store(w, static_cast<uint32_t *>(byte_at(p, off)));
    }
};

// Probe payload: 1 + 32*(global x-fastest node index) + q, carried as float bits.
__device__ __forceinline__ float probe_id(int x, int y, int z, int nx, int ny, int q) {
    const uint32_t s = uint32_t(x) + uint32_t(nx) * (uint32_t(y) + uint32_t(ny) * uint32_t(z));
    return __uint_as_float(1u + 32u * s + uint32_t(q));
}

template <int TX, int TY>
struct Geo {
    static constexpr int NT = TX * TY;
    static constexpr int WP = TX + 2;            // LDS row pitch (with ring)
    static constexpr int P = (TX + 2) * (TY + 2);  // LDS plane size
};

// Ring-fill jobs: every ring cell population that points into the footprint.
// Sides: 0 W (ring x=-1, cx=+1 pops, from the east faces of column bx-1),
//        1 E (ring x=TX, cx=-1, west faces of column bx+1),
//        2 S (ring y=-1, cy=+1, north faces of row by-1; corners excluded),
//        3 N (ring y=TY, cy=-1, south faces of row by+1; corners excluded).
template <int Q, int TX, int TY>
struct Ring {
    static constexpr int QF = VS<Q>::QF;
    __host__ __device__ static constexpr int side_count(int side, int k) {
        if (side < 2) return TY;
        const int q = face_q<Q>(1, side == 2 ? 1 : -1, k);
        return TX - absi(cxq(q));
    }
    __host__ __device__ static constexpr int total() {
        int t = 0;
        for (int s = 0; s < 4; s++)
            for (int k = 0; k < QF; k++) t += side_count(s, k);
        return t;
    }
    static constexpr int NJ = total();
    static constexpr int JPT = (NJ + TX * TY - 1) / (TX * TY);
};

// LDS base of the slot holding up-going population q in bank 0 or 1 (bank 1 is
// bank 0 shifted by NU slots, see lds_slot).
template <int Q, int TX, int TY>
__host__ __device__ constexpr int u_addr(int q, int bank) {
    return (lds_slot<Q>(q) + bank * VS<Q>::NU) * Geo<TX, TY>::P;
}
static_assert(lds_slot<19>(9) == 12 && lds_slot<19>(18) == 15 && lds_slot<19>(17) == 11, "D3Q19 LDS slots");
static_assert(lds_slot<27>(9) == 16 && lds_slot<27>(25) == 23 && lds_slot<27>(26) == 15, "D3Q27 LDS slots");
// Ring job in three registers: LDS word offsets of bank 0 / bank 1 (16 bits
// each; `lds` = -1 when the thread has no job), the face-set offset without the
// plane term (-1 when outside the domain), and plane stride * 4 + (dz + 1).
struct Job {
    int lds;  // bank0 | bank1 << 16 (16-bit LDS word addresses, so the sign bit may be set), or -1: no job
    int gb;
    int zsd;
    __device__ __forceinline__ bool valid() const { return lds != -1; }
    __device__ __forceinline__ int lds0() const { return lds & 0xFFFF; }
    __device__ __forceinline__ int lds1() const { return (lds >> 16) & 0xFFFF; }
    __device__ __forceinline__ int zs() const { return zsd >> 2; }
    __device__ __forceinline__ int dz() const { return (zsd & 3) - 1; }
};

// Offsets of one population inside the face blocks.  X faces [bx][slot][z][side][y],
// Y faces [by][slot][z][side][x]: the two sides of a column's face (W/E, S/N) sit
// next to each other, so a face row's lanes of both sides differ only in a
// small lane offset (side * ny + y) from a base that is uniform in the workgroup.
// Z faces [side][chunk][slot][y][x].
// All offsets are 32-bit: the host rejects meshes whose arrays exceed 2^31 words.
template <int Q>
__host__ __device__ __forceinline__ int xf_off(const StepParams &p, int s, int bx, int k, int z, int y) {
    return (((bx * VS<Q>::QF + k) * p.nz + z) * 2 + s) * p.ny + y;
}
template <int Q>
__host__ __device__ __forceinline__ int yf_off(const StepParams &p, int s, int by, int k, int z, int x) {
    return p.yoff + ((((by * VS<Q>::QF + k) * p.nz + z) * 2 + s) * p.nx + x);
}
template <int Q>
__device__ __forceinline__ int zf_off(const StepParams &p, int s, int c, int k, int y, int x) {
    return p.zoff + (((s * p.nc + c) * VS<Q>::QF + k) * p.ny + y) * p.nx + x;
}
// Ring job j of this workgroup: (side, face slot k, position) decoded against the
// compile-time segment table, so the population index is a constant per branch.
template <int Q, int TX, int TY>
__device__ __forceinline__ Job make_job(int j, const StepParams &p, int bx, int by) {
    using G = Geo<TX, TY>;
    constexpr int QF = VS<Q>::QF;
    // both banks' word addresses below 0xFFFF: a packed pair never reads as -1
    static_assert((VS<Q>::NSLOT + VS<Q>::NU) * G::P < 0xFFFF, "ring-job LDS addresses fit 16 bits");
    Job job{-1, -1, 0};
    int start = 0;
    unroll<0, 4>([&](auto sc) {
        constexpr int side = decltype(sc)::value;
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int axis = side < 2 ? 0 : 1;
            constexpr int sign = (side == 0 || side == 2) ? 1 : -1;
            constexpr int q = face_q<Q>(axis, sign, k);
            constexpr int cnt = Ring<Q, TX, TY>::side_count(side, k);
            const int pos = j - start;
            if (pos >= 0 && pos < cnt) {
                int rx, ry;
                if constexpr (side < 2) {
                    rx = side == 0 ? -1 : TX;
                    ry = pos - cyq(q);  // ry in [-cy, TY-1-cy]
                } else {
                    ry = side == 2 ? -1 : TY;
                    rx = pos + (cxq(q) < 0 ? 1 : 0);  // rx in [max(0,-cx), min(TX-1, TX-1-cx)]
                }
                const int cell = (ry + 1) * G::WP + (rx + 1);
                const int l0 = lds_slot<Q>(q) * G::P + cell;
                const int l1 = czq(q) == 1 ? u_addr<Q, TX, TY>(q, 1) + cell : l0;
                job.lds = l0 | (l1 << 16);
                constexpr int dzq = -czq(q);
                const int gx = bx * TX + rx, gy = by * TY + ry;
                const bool inside = gx >= 0 && gx < p.nx && gy >= 0 && gy < p.ny;
                if constexpr (side < 2) {
                    const int col = side == 0 ? bx - 1 : bx + 1;
                    constexpr int s_face = side == 0 ? 1 : 0;  // east (1) / west (0) faces
                    job.zsd = 2 * p.ny * 4 + (dzq + 1);
                    job.gb = inside ? xf_off<Q>(p, s_face, col, k, 0, gy) : -1;
                } else {
                    const int row = side == 2 ? by - 1 : by + 1;
                    constexpr int s_face = side == 2 ? 1 : 0;
                    job.zsd = 2 * p.nx * 4 + (dzq + 1);
                    job.gb = inside ? yf_off<Q>(p, s_face, row, k, 0, gx) : -1;
                }
            }
            start += cnt;
        });
    });
    return job;
}

// Raw face word for destination plane zdest (address clamped into the buffer so
// the load is unconditional); job_value() applies the validity afterwards.
__device__ __forceinline__ float job_raw(const Job &j, const float *__restrict__ fin, int zdest, int nz) {
    int zs = zdest + j.dz();
    zs = zs < 0 ? 0 : (zs >= nz ? nz - 1 : zs);
    const int off = j.gb < 0 ? 0 : j.gb + zs * j.zs();
    return fin[off];
}
__device__ __forceinline__ float job_value(const Job &j, float raw, int zdest, int nz) {
    const int zs = zdest + j.dz();
    return (j.gb >= 0 && zs >= 0 && zs < nz) ? raw : 0.f;
}

// Population q of node (x,y,z) from its rebuild terms (or its probe id).
template <int Q, int q, bool PROBE>
__device__ __forceinline__ float pop_of(const Rec &r, int x, int y, int z,
                                        const StepParams &p) {
    if constexpr (PROBE)
        return probe_id(x, y, z, p.nx, p.ny, q);
    else
        return rec_dev<Q, q>(r);
}

// Outgoing x/y face populations of one node after collision (or its probe ids).
template <int Q, int TX, int TY, bool PROBE>
__device__ __forceinline__ void write_x_faces(const StepParams &p, float *__restrict__ fout, const Rec &r, int tx,
                                              int x, int y, int z, int bx) {
    constexpr int QF = VS<Q>::QF;
    if (tx == 0 && bx > 0)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(0, -1, k);
            fout[xf_off<Q>(p, 0, bx, k, z, y)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
    if (tx == TX - 1 && bx < p.nbx - 1)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(0, 1, k);
            fout[xf_off<Q>(p, 1, bx, k, z, y)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
}
template <int Q, int TX, int TY, bool PROBE>
__device__ __forceinline__ void write_y_faces(const StepParams &p, float *__restrict__ fout, const Rec &r, int ty,
                                              int x, int y, int z, int by) {
    constexpr int QF = VS<Q>::QF;
    if (ty == 0 && by > 0)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(1, -1, k);
            fout[yf_off<Q>(p, 0, by, k, z, x)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
    if (ty == TY - 1 && by < p.nby - 1)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(1, 1, k);
            fout[yf_off<Q>(p, 1, by, k, z, x)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
}
template <int Q, int TX, int TY, bool PROBE>
__device__ __forceinline__ void write_xy_faces(const StepParams &p, float *__restrict__ fout, const Rec &r, int tx,
                                               int ty, int x, int y, int z, int bx, int by) {
    write_x_faces<Q, TX, TY, PROBE>(p, fout, r, tx, x, y, z, bx);
    write_y_faces<Q, TX, TY, PROBE>(p, fout, r, ty, x, y, z, by);
}

// Outgoing z face populations of a node on the chunk's first plane (down-going,
// read by chunk c-1) or last plane (up-going, read by chunk c+1).
template <int Q, bool PROBE>
__device__ __forceinline__ void write_z_faces(const StepParams &p, float *__restrict__ fout, const Rec &r, int x,
                                              int y, int z, int c, int z0, int z1) {
    constexpr int QF = VS<Q>::QF;
    if (z == z0 && c > 0)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(2, -1, k);
            fout[zf_off<Q>(p, 0, c, k, y, x)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
    if (z == z1 && c < p.nc - 1)
        unroll<0, QF>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            constexpr int q = face_q<Q>(2, 1, k);
            fout[zf_off<Q>(p, 1, c, k, y, x)] = pop_of<Q, q, PROBE>(r, x, y, z, p);
        });
}

// Edge nodes of the footprint (the ones that own x/y face populations),
// numbered bottom row, top row, left column, right column (corners once).
template <int TX, int TY>
struct Edge {
    // entries: the S row (ty = 0) and the N row (ty = TY-1), then the W column
    // (tx = 0) and the E column (tx = TX-1), corners in both lists, so that a
    // row entry writes only y faces and a column entry only x faces
    static constexpr int NE = 2 * TX + 2 * TY;
    __device__ static __forceinline__ int row(int tx, int ty) {
        return ty == 0 ? tx : (ty == TY - 1 ? TX + tx : -1);
    }
    __device__ static __forceinline__ int col(int tx, int ty) {
        return tx == 0 ? 2 * TX + ty : (tx == TX - 1 ? 2 * TX + TY + ty : -1);
    }
    __device__ static __forceinline__ void coords(int e, int &tx, int &ty) {
        if (e < TX) { tx = e; ty = 0; }
        else if (e < 2 * TX) { tx = e - TX; ty = TY - 1; }
        else if (e < 2 * TX + TY) { tx = 0; ty = e - 2 * TX; }
        else { tx = TX - 1; ty = e - 2 * TX - TY; }
    }
};

// Population-rebuild terms from moments as stored: the raw words of
// load_plane_raw.  FP32: rec_setup.  FP16: h[j] (j < 5) holds node word j
// (moment 2j in its low half, 2j + 1 in its high half), decoded by
// mixed-precision FMAs, FP16 operand taken directly (no separate conversion);
// the decode factors are equal within the groups {1,2,3}, {4,7,9}, {5,6,8} (the
// codec's u, Pi_aa and Pi_ab scales) and the density code is 3 (rho - 1) itself.
// Every population the step streams or a face buffer holds is rebuilt by this
// one function, so all readers of a stored state see the same bits.
template <typename ST>
__device__ __forceinline__ Rec rec_setup_st(const float (&h)[10]) {
    if constexpr (sizeof(ST) == 4) {
        return rec_setup(h);
    } else {
        auto hv = [&](int k) { return (k & 1) ? Store<ST>::from_raw_hi(h[k / 2]) : Store<ST>::from_raw(h[k / 2]); };
        Rec r;
        const float rho = fmaf(hv(0), codec_dec(0), 1.0f);
        float sd = fmaf(hv(9), codec_dec(4), 0.0f);
        sd = fmaf(hv(7), codec_dec(4), sd);
        sd = fmaf(hv(4), codec_dec(4), sd);
        r.a = fmaf(rho, -sd, hv(0));  // 3 rho' - rho (mxx + myy + mzz)
        const float t = 3.0f * rho;
        const float tu = t * codec_dec(1), td = t * codec_dec(4), to = t * codec_dec(5);
        r.M[1] = fmaf(hv(1), tu, 0.0f);
        r.M[2] = fmaf(hv(2), tu, 0.0f);
        r.M[3] = fmaf(hv(3), tu, 0.0f);
        r.M[4] = fmaf(hv(4), td, 0.0f);
        r.M[7] = fmaf(hv(7), td, 0.0f);
        r.M[9] = fmaf(hv(9), td, 0.0f);
        r.M[5] = fmaf(hv(5), to, 0.0f);
        r.M[6] = fmaf(hv(6), to, 0.0f);
        r.M[8] = fmaf(hv(8), to, 0.0f);
        r.M[0] = 0.f;
        return r;
    }
}

// Edge-node moments as stored, parked in LDS for the deferred face flush
// (park: [10][NE] FP32 values, or [5][NE] FP16 node words) and read back in
// the raw-word form of load_plane_raw.
template <typename ST, int NE>
__device__ __forceinline__ void park_read(const ST *park, int e, float (&m)[10]) {
    if constexpr (sizeof(ST) == 4) {
#pragma unroll
        for (int k = 0; k < 10; k++) m[k] = park[k * NE + e];
    } else {
        const uint32_t *pw = reinterpret_cast<const uint32_t *>(park);
#pragma unroll
        for (int j = 0; j < 5; j++) m[j] = __uint_as_float(pw[j * NE + e]);
    }
}

// x/y face populations of one finished plane, written cooperatively: edge
// nodes park their post-collision moments in LDS (edge_m, 10 x NE values as
// stored: FP32, or 5 x NE FP16 node words that rec_setup_st decodes) and
// the first NE threads -- whole waves where TX is a multiple of 32 -- rebuild
// and store the face populations, instead of every wave rebuilding them for
// its few edge lanes.  Row entries write the y faces, column entries the x
// faces.  The two faces of one kind carry opposite populations (slot k of the
// +axis face holds q, slot k of the -axis face opposite(q)), which share the
// even part of the rebuild: a lane forms its face's population with its own
// sign (rec_signed), so the S and N rows (W and E columns) run one instruction
// stream, and the store address is a uniform base per slot plus a 32-bit lane
// offset.
template <int Q, int TX, int TY, bool PROBE, typename ST>
__device__ __forceinline__ void flush_xy_faces(const StepParams &p, float *__restrict__ fout,
                                               const ST *edge_m, int tid, int bx, int by, int z) {
    using E = Edge<TX, TY>;
    constexpr int QF = VS<Q>::QF;
    if (tid >= E::NE) return;
    // the plane index is made opaque here: otherwise the compiler keeps every
    // face slot's write offset as a loop induction variable in its own VGPR,
    // advanced on every plane by every wave, although only the waves holding
    // edge nodes ever store through them (18 VALU per plane and wave for D3Q19)
    asm volatile("" : "+s"(z));
    const bool row = tid < 2 * TX;
    // side 0: the S row / W column (populations with c = -1), side 1: N / E (c = +1)
    const int side = row ? (tid >= TX ? 1 : 0) : (tid >= 2 * TX + TY ? 1 : 0);
    const bool neighbour = row ? (side ? by < p.nby - 1 : by > 0) : (side ? bx < p.nbx - 1 : bx > 0);
    if (!neighbour) return;
    int tx, ty;
    E::coords(tid, tx, ty);
    float m[10];
    park_read<ST, E::NE>(edge_m, tid, m);
    const Rec r = rec_setup_st<ST>(m);
    const float sgn = side ? 1.f : -1.f;
    const int x = bx * TX + tx, y = by * TY + ty;
    auto put = [&](auto kc, auto axc) {
        constexpr int k = decltype(kc)::value, axis = decltype(axc)::value;
        constexpr int q = face_q<Q>(axis, 1, k);
        static_assert(face_q<Q>(axis, -1, k) == opposite(q), "face slots pair opposite populations");
        float v;
        if constexpr (PROBE)
            v = probe_id(x, y, z, p.nx, p.ny, side ? q : opposite(q));
        else
            v = rec_signed<Q, q>(r, sgn);
        float *base = fout + (axis == 0 ? xf_off<Q>(p, 0, bx, k, z, 0) : yf_off<Q>(p, 0, by, k, z, 0));
        const unsigned off = (unsigned)(axis == 0 ? side * p.ny + y : side * p.nx + x) * 4u;
        *static_cast<float *>(byte_at(base, off)) = v;
    };
    if (row)
        unroll<0, QF>([&](auto kc) { put(kc, std::integral_constant<int, 1>{}); });
    else
        unroll<0, QF>([&](auto kc) { put(kc, std::integral_constant<int, 0>{}); });
}

// Plane block of this workgroup's tile in plane z (uniform).
__device__ __forceinline__ unsigned plane_block(const StepParams &p, int bx, int by, int z) {
    return (unsigned)((z * p.nby + by) * p.nbx + bx);
}
// First node of the block in the internal node order (the FP64 means, k, S
// and the residual fields are indexed by it); add the thread index.
template <int TX, int TY>
__device__ __forceinline__ unsigned plane_base(const StepParams &p, int bx, int by, int z) {
    return plane_block(p, bx, by, z) * (unsigned)(TX * TY);
}
// Byte offset of the block's first node in every moment array.
template <typename ST>
__device__ __forceinline__ unsigned mom_plane_off(const StepParams &p, int bx, int by, int z) {
    return plane_block(p, bx, by, z) * (unsigned)p.mplane * Store<ST>::kBytes;
}
// Byte offset of internal node i in every moment array.
template <typename ST>
__device__ __forceinline__ unsigned mom_off(const StepParams &p, int64_t i) {
    const unsigned u = (unsigned)i;
    return ((u >> p.mshift) * (unsigned)p.mplane + (u & ((1u << p.mshift) - 1u))) * Store<ST>::kBytes;
}

// Raw words of one stored plane: 10 FP32 moments, or 5 FP16 node words.
template <typename ST>
constexpr int kRawWords = sizeof(ST) == 4 ? 10 : 5;

// The same plane as stored, in m[0 .. kRawWords) (FP16: the five node words,
// not yet scaled by the codec's decode factors, which rec_setup_st applies).
// PROBE: zeros.
template <typename ST, int TX, int TY, bool PROBE = false>
__device__ __forceinline__ void load_plane_raw(const StepParams &p, int bx, int by, int z, int tid,
                                               float (&m)[10]) {
    const unsigned off = mom_plane_off<ST>(p, bx, by, z) + (unsigned)tid * Store<ST>::kBytes;
    if constexpr (PROBE) {
#pragma unroll
        for (int k = 0; k < 10; k++) m[k] = 0.f;
    } else if constexpr (sizeof(ST) == 4) {
#pragma unroll
        for (int k = 0; k < 10; k++) m[k] = Store<ST>::load_raw(p.mom[k], off);
    } else {
#pragma unroll
        for (int j = 0; j < 5; j++) m[j] = Store<ST>::load_word(p.mom[2 * j], off);
    }
}

// FP16 storage rounds the stored 3u stochastically (v_cvt_sr_f16_f32: up with
// probability equal to the remainder's fraction of an ulp): round-to-nearest
// of a slowly changing field errs the same way step after step and the error
// grows like a bias, while stochastic rounding is unbiased.  The random word of
// node (x, y, z) at step t is seed(x, y, t) + z * 0x9E3779B9 (a Weyl sequence
// down the column; seed = lowbias32 hash, C. Wellons, of the column index and
// of t), rotated by 0/11/22 bits for u, v, w: reproducible, independent of the
// tiling, and one add per node in the plane loop.
__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
// v rounded stochastically into half `hi` of w (the other half kept)
template <bool HI>
__device__ __forceinline__ half2_t sr_into(half2_t w, float v, unsigned r) {
    return __builtin_amdgcn_cvt_sr_f16_f32(w, v, r, HI);
}
// Function-object values of one node from its stored (scaled) moments:
// k = kineticEnergy.cuh:71-85; S = strainRateTensor.cuh:71-84, i.e.
// S_aa = 3(m_a^2 - m_aa)/(2 * 4.5 tau) = (m_a^2 - m_aa) s_diag and
// S_ab = 3(m_a m_b - m_ab)/(2 * 9 tau) = (m_a m_b - m_ab) s_off with the
// coefficients 1/(3 tau), 1/(6 tau) formed once on the host.
__device__ __forceinline__ void fo_values(const float m[10], float s_diag, float s_off, float &kk, float S[6]) {
    kk = 0.5f * sqrtf(m[1] * m[1] + m[2] * m[2] + m[3] * m[3]);
    S[0] = (m[1] * m[1] - m[4]) * s_diag;
    S[1] = (m[1] * m[2] - m[5]) * s_off;
    S[2] = (m[1] * m[3] - m[6]) * s_off;
    S[3] = (m[2] * m[2] - m[7]) * s_diag;
    S[4] = (m[2] * m[3] - m[8]) * s_off;
    S[5] = (m[3] * m[3] - m[9]) * s_diag;
}

// The fused step (body in step_body.inc).  PROBE replaces the moment algebra by
// integer ids so that tests can read back the exact source of every population.
//
// One cavity: workgroup b owns column (bx, by) and chunk c of the domain.  The
// parameters are the kernel argument itself.
template <int Q, int TX, int TY, typename ST, bool FO, bool PROBE>
__global__ __launch_bounds__(TX *TY, Q == 19 ? 4 : 3) void step_kernel(const StepParams p) {
    const int tid = threadIdx.x;
    if constexpr (FO || sizeof(ST) == 2) {
        if (blockIdx.x == 0 && tid == 0 && p.clock_next) {
            const int64_t t = p.clock->t + 1;
            p.clock_next->t = t;
            p.clock_next->inv = 1.0 / double(t + 1);
        }
    }
#define LBM_WORK_ITEM                      \
    const int b = blockIdx.x;              \
    const int bx = b % p.nbx;              \
    const int by = (b / p.nbx) % p.nby;    \
    const int c = b / (p.nbx * p.nby);
#include "step_body.inc"
#undef LBM_WORK_ITEM
}

// Many cavities in one launch: work item i of a tile-to-case table names a case
// and one (column, chunk) of it; the case's parameters and buffers come from a
// per-case table in device memory (read with scalar loads).  One launch
// advances every listed case by one step (FP32 storage, no function objects).
// PERSIST = false: one workgroup per item.  PERSIST = true: a persistent grid
// whose workgroups stride over the items; it is built for one workgroup fewer
// per CU, because the item loop needs more registers than a single pass.
struct BatchItem {
    int cs, bx, by, c;
};
template <int Q, int TX, int TY, bool PERSIST>
__global__ __launch_bounds__(TX *TY, PERSIST ? (Q == 19 ? 3 : 2) : (Q == 19 ? 4 : 3)) void batch_step_kernel(
    const StepParams *__restrict__ cases, const BatchItem *__restrict__ items, int count) {
    using ST = float;
    constexpr bool FO = false, PROBE = false;
    if constexpr (!PERSIST) {
        if ((int)blockIdx.x >= count) return;
        const BatchItem it = items[blockIdx.x];
        const StepParams &p = cases[it.cs];
        const int tid = threadIdx.x;
#define LBM_WORK_ITEM const int bx = it.bx, by = it.by, c = it.c;
#include "step_body.inc"
    } else {
        for (int i = blockIdx.x; i < count; i += gridDim.x) {
            const BatchItem it = items[i];
            const StepParams &p = cases[it.cs];
            // an opaque copy of the thread index per item keeps per-thread
            // addressing from being hoisted out of the item loop (it would stay
            // live through the whole body and spill)
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            {
#include "step_body.inc"
            }
            __syncthreads();  // the next item rewrites the LDS the face flush above read
        }
    }
#undef LBM_WORK_ITEM
}

// Face buffers from the current moments (post-collision populations of every
// face node): the device-side replacement of haloFace::initialise_pop
// (haloFace.cuh:244-449).  Same grid as the step, and the same rebuild from the
// moments as stored (rec_setup_st) as the step's own face writes, so a restart
// reproduces the faces an uninterrupted run would hold bit for bit.
template <int Q, int TX, int TY, typename ST, bool PROBE>
__global__ __launch_bounds__(TX *TY) void fill_faces_kernel(const StepParams p) {
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int b = blockIdx.x;
    const int bx = b % p.nbx;
    const int by = (b / p.nbx) % p.nby;
    const int c = b / (p.nbx * p.nby);
    const int x = bx * TX + tx, y = by * TY + ty;
    const int z0 = c * p.lz, z1 = z0 + p.lz - 1;
    for (int z = z0; z <= z1; z++) {
        float m[10];
        if constexpr (!PROBE)
            load_plane_raw<ST, TX, TY>(p, bx, by, z, tid, m);
        else
            for (int k = 0; k < 10; k++) m[k] = 0.f;
        const Rec r = rec_setup_st<ST>(m);
        write_xy_faces<Q, TX, TY, PROBE>(p, p.fout, r, tx, ty, x, y, z, bx, by);
        write_z_faces<Q, PROBE>(p, p.fout, r, x, y, z, c, z0, z1);
    }
}

}  // namespace lbm
#pragma float_control(pop)
