// lbm_capi.hip -- the C ABI of include/lbm_amd.h: handle lifecycle, layout
// conversion, stepping through hipGraphs, fused function objects, FP64
// reductions and the integer streaming probe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lbm_amd.h"
#include "aux_kernels.hpp"
#include "lbm_internal.hpp"

using namespace lbm;

namespace {

thread_local std::string g_create_error;

// ---------------------------------------------- scratch of lbm_reduce_noalloc
// The per-block FP64 partial sums of the residual reduction (3 per block) and
// the three totals, in device memory that belongs to the library's code object
// rather than to any allocation: one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself (in host code the name is a stand-in,
// not a device address); it uses the address hipGetSymbolAddress returns for
// the current device, looked up once per device by scratch_on().
constexpr int kScratchBlocks = 1024;  // reduction blocks whose partials fit
constexpr int kScratchWords = 3 * kScratchBlocks + 3;
static __device__ __align__(256) double scratch[kScratchWords];

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;

// Device address of `scratch` on `device` (which must be current).
double *scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    double *ptr = nullptr;
    HIPX(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipError{hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)"};
    g_scratch_ptr[device] = ptr;
    return ptr;
}

// ------------------------------------------------------------ dispatch
template <int Q, int TX, int TY>
void launch_step_t(const StepParams &p, int grid, hipStream_t s, int storage, bool fo, bool probe) {
    const dim3 g(grid), b(TX * TY);
    if (probe)
        hipLaunchKernelGGL((step_kernel<Q, TX, TY, float, false, true>), g, b, 0, s, p);
    else if (storage == LBM_STORE_FP32) {
        if (fo)
            hipLaunchKernelGGL((step_kernel<Q, TX, TY, float, true, false>), g, b, 0, s, p);
        else
            hipLaunchKernelGGL((step_kernel<Q, TX, TY, float, false, false>), g, b, 0, s, p);
    } else {
        if (fo)
            hipLaunchKernelGGL((step_kernel<Q, TX, TY, _Float16, true, false>), g, b, 0, s, p);
        else
            hipLaunchKernelGGL((step_kernel<Q, TX, TY, _Float16, false, false>), g, b, 0, s, p);
    }
}

template <int Q, int TX, int TY>
void launch_fill_t(const StepParams &p, int grid, hipStream_t s, int storage, bool probe) {
    const dim3 g(grid), b(TX * TY);
    if (probe)
        hipLaunchKernelGGL((fill_faces_kernel<Q, TX, TY, float, true>), g, b, 0, s, p);
    else if (storage == LBM_STORE_FP32)
        hipLaunchKernelGGL((fill_faces_kernel<Q, TX, TY, float, false>), g, b, 0, s, p);
    else
        hipLaunchKernelGGL((fill_faces_kernel<Q, TX, TY, _Float16, false>), g, b, 0, s, p);
}

template <int Q>
bool launch_step_q(int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool fo,
                   bool probe) {
    if (tx == 16 && ty == 16) launch_step_t<Q, 16, 16>(p, grid, s, storage, fo, probe);
    else if (tx == 32 && ty == 8) launch_step_t<Q, 32, 8>(p, grid, s, storage, fo, probe);
    else if (tx == 8 && ty == 8) launch_step_t<Q, 8, 8>(p, grid, s, storage, fo, probe);
    else if (tx == 32 && ty == 16) launch_step_t<Q, 32, 16>(p, grid, s, storage, fo, probe);
    else if (tx == 32 && ty == 32) launch_step_t<Q, 32, 32>(p, grid, s, storage, fo, probe);
    else if constexpr (Q == 19) {
        if (tx == 64 && ty == 16) launch_step_t<Q, 64, 16>(p, grid, s, storage, fo, probe);
        else return false;
    } else return false;
    return true;
}
template <int Q>
bool launch_fill_q(int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool probe) {
    if (tx == 16 && ty == 16) launch_fill_t<Q, 16, 16>(p, grid, s, storage, probe);
    else if (tx == 32 && ty == 8) launch_fill_t<Q, 32, 8>(p, grid, s, storage, probe);
    else if (tx == 8 && ty == 8) launch_fill_t<Q, 8, 8>(p, grid, s, storage, probe);
    else if (tx == 32 && ty == 16) launch_fill_t<Q, 32, 16>(p, grid, s, storage, probe);
    else if (tx == 32 && ty == 32) launch_fill_t<Q, 32, 32>(p, grid, s, storage, probe);
    else if constexpr (Q == 19) {
        if (tx == 64 && ty == 16) launch_fill_t<Q, 64, 16>(p, grid, s, storage, probe);
        else return false;
    } else return false;
    return true;
}

// Workgroups of one step-kernel instantiation resident per CU (0 if the runtime
// cannot tell).
template <int Q, int TX, int TY>
int step_occupancy_t(int storage, bool fo) {
    int nb = 0;
    hipError_t e;
    if (storage == LBM_STORE_FP32)
        e = fo ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<Q, TX, TY, float, true, false>, TX * TY, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<Q, TX, TY, float, false, false>, TX * TY, 0);
    else
        e = fo ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<Q, TX, TY, _Float16, true, false>, TX * TY, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_kernel<Q, TX, TY, _Float16, false, false>, TX * TY, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}
template <int Q>
int step_occupancy_q(int tx, int ty, int storage, bool fo) {
    if (tx == 16 && ty == 16) return step_occupancy_t<Q, 16, 16>(storage, fo);
    if (tx == 32 && ty == 8) return step_occupancy_t<Q, 32, 8>(storage, fo);
    if (tx == 8 && ty == 8) return step_occupancy_t<Q, 8, 8>(storage, fo);
    if (tx == 32 && ty == 16) return step_occupancy_t<Q, 32, 16>(storage, fo);
    if (tx == 32 && ty == 32) return step_occupancy_t<Q, 32, 32>(storage, fo);
    if constexpr (Q == 19) {
        if (tx == 64 && ty == 16) return step_occupancy_t<Q, 64, 16>(storage, fo);
    }
    return 0;
}

}  // namespace

namespace lbm {

bool tile_supported(int Q, int tx, int ty) {
    return (tx == 16 && ty == 16) || (tx == 32 && ty == 8) || (tx == 8 && ty == 8) || (tx == 32 && ty == 16) ||
           (tx == 32 && ty == 32) || (Q == 19 && tx == 64 && ty == 16);
}

void launch_step(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool fo,
                 bool probe) {
    const bool ok = Q == 19 ? launch_step_q<19>(tx, ty, p, grid, s, storage, fo, probe)
                            : launch_step_q<27>(tx, ty, p, grid, s, storage, fo, probe);
    if (!ok) throw ArgError{LBM_ERR_ARG, "unsupported tile"};
    HIPX(hipGetLastError());
}

void launch_fill(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool probe) {
    const bool ok = Q == 19 ? launch_fill_q<19>(tx, ty, p, grid, s, storage, probe)
                            : launch_fill_q<27>(tx, ty, p, grid, s, storage, probe);
    if (!ok) throw ArgError{LBM_ERR_ARG, "unsupported tile"};
    HIPX(hipGetLastError());
}

int64_t face_words(int Q, int nx, int ny, int nz, int tx, int ty, int lz, int64_t *yoff, int64_t *zoff) {
    const int QF = Q == 19 ? 5 : 9;
    const int64_t nbx = nx / tx, nby = ny / ty, nc = nz / lz;
    const int64_t xw = 2 * nbx * QF * (int64_t)nz * ny;
    const int64_t yw = 2 * nby * QF * (int64_t)nz * nx;
    const int64_t zw = 2 * nc * QF * (int64_t)ny * nx;
    if (yoff) *yoff = xw;
    if (zoff) *zoff = xw + yw;
    return xw + yw + zw;
}

// Planes per work item.  A launch runs in rounds of `slots` resident
// workgroups; a work item of lz planes costs about (lz + h) plane times (h:
// pipeline fill and drain, ring and z-face set-up; 4 for D3Q19, 8 for D3Q27
// with its nine face slots per side) at bpn(lz) bytes per
// node, bpn = M + 16 QF (1/TX + 1/TY + 1/lz) with M = 80 moment bytes (40 in
// FP16) and the face words.  The choice minimises rounds x (lz + 4) x bpn(lz)
// over the divisors of nz in [4, 64]; with FP32 storage a single round costs
// 10% more (no other workgroup overlaps the first and last planes of the
// HBM-bound step), with FP16 it does not (its step is instruction-bound).
// Measured on MI355X (MLUPS, best choice = the model's): D3Q19 32x16 at 128^3
// lz 8 (45.0k; lz 4: 40.9k, 16: 40.5k), 192^3 lz 16 (52.8k; 24: 43.7k), 256^3
// lz 32 (63.0k; 64: 61.6k), FP16 256^3 lz 64 (85.4k; 32: 81.9k); D3Q27 16x16
// at 128^3 lz 16 (31.4k; 8: 27.7k), 192^3 lz 12 (40.0k; 16: 37.8k), 256^3 lz
// 32 (42.6k; 64: 38.4k), 384^3 lz 48 (45.1k; 64: 44.0k); D3Q27 FP16 32x8 at
// 256^3 lz 32 (49.1k; 16: 46.6k, 64: 42.1k), 384^3 lz 48 (53.7k; 32: 52.7k).
// With the FP64 means on (mean_bytes per node read + written), a single round
// pays no penalty either: D3Q19 32x32 at 256^3 lz 64 (616 us; 32: 636, 16: 649).
// D3Q27 32x16 (two workgroups per CU): 256^3 FP32 lz 64 (266.1 us; 32: 272.7),
// FP16 lz 64 (180.0; 32: 182.8); 384^3 FP32 lz 48 (1054.9; 64: 1059.0, 32:
// 1081.9), FP16 lz 48 (631.7; 64: 640.0, 32: 664.9) -- h = 4 would pick 32.
int choose_lz(int Q, int nx, int ny, int nz, int tx, int ty, int64_t slots, bool fp16, int mean_bytes) {
    const int64_t cols = (int64_t)(nx / tx) * (ny / ty);
    const double face = 16.0 * (Q == 19 ? 5 : 9);
    const double mom = (fp16 ? 40.0 : 80.0) + mean_bytes;
    const bool single_round_penalty = !fp16 && mean_bytes == 0;
    int best = -1;
    double best_cost = 0.0;
    for (int lz = 4; lz <= std::min(nz, 64); lz++) {
        if (nz % lz) continue;
        const int64_t items = cols * (nz / lz);
        const int64_t rounds = (items + slots - 1) / slots;
        const double bpn = mom + face * (1.0 / tx + 1.0 / ty + 1.0 / lz);
        const double cost =
            (double)rounds * (lz + (Q == 19 ? 4 : 8)) * bpn * (rounds == 1 && single_round_penalty ? 1.1 : 1.0);
        if (best < 0 || cost < best_cost) {
            best = lz;
            best_cost = cost;
        }
    }
    if (best > 0) return best;
    for (int lz = std::min(nz, 64); lz >= 1; lz--)  // nz has no divisor in [4, 64]
        if (nz % lz == 0) return lz;
    return nz;
}

// Workgroups resident at once: CUs x workgroups per CU of the step kernel the
// case will run (registers and LDS of that instantiation, from the runtime;
// failing that, 4 waves per SIMD for D3Q19 and 3 for D3Q27, 4 SIMDs per CU).
int64_t resident_workgroups(int Q, int tx, int ty, int device, int storage, bool fo) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 256;  // MI355X
    if (hipSetDevice(device) != hipSuccess) (void)hipGetLastError();  // the occupancy query asks the current device
    int per_cu = Q == 19 ? step_occupancy_q<19>(tx, ty, storage, fo) : step_occupancy_q<27>(tx, ty, storage, fo);
    if (per_cu <= 0) {
        const int waves = (tx * ty + 63) / 64;
        per_cu = std::max(1, (4 * (Q == 19 ? 4 : 3)) / waves);
    }
    return (int64_t)cus * per_cu;
}

void choose_tile(int Q, int nx, int ny, bool fp16, bool means, int *tx, int *ty) {
    // measured on MI355X at 256^3: the D3Q19 FP32 step is fastest with 32x16
    // footprints (8 waves, two workgroups per CU): 246.3 us against 249.8 on
    // 32x32x64, 255.3 on 64x16x64.  With FP16 storage or the FP64 means on, 32x32
    // (one 16-wave workgroup per CU, 6.25 instead of 10 face bytes per node) wins:
    // FP16 132.9 vs 138.6 us (32x16x64), with the means FP32 616 vs 661 us and
    // FP16 554 vs 568 us.
    // D3Q27 without the means (one up-going LDS bank) with FP32 storage runs
    // 32x16 (two workgroups per CU) unless its work items span more than one
    // round of resident workgroups with a partial last one, then 32x32
    // (auto_tiling): 256^3 265 us on 32x16x64 (512 items, one round) vs 272 us on
    // 32x32x64 and 350 on 16x16; 384^3 996 us on 32x32x32 vs 1055 on 32x16x48
    // (4.5 rounds) and 1183 on 16x16.  With FP16
    // storage the instruction-bound step gains from 32x32 (one workgroup per CU,
    // fewer edge nodes to flush per node): 256^3 170.4 vs 180.7 us (32x16),
    // 241 us (32x8); 384^3 628 vs 632 us.  With the means (two banks) FP16 runs
    // 9% faster at 32x8 than at 16x16 (256^3: 342 vs 373 us; 384^3: 1054 vs 1152 us).
    if (Q == 19 && (means || fp16) && nx % 32 == 0 && ny % 32 == 0) { *tx = 32; *ty = 32; }
    else if (Q == 19 && nx % 32 == 0 && ny % 16 == 0) { *tx = 32; *ty = 16; }
    else if (Q == 27 && !means && fp16 && nx % 32 == 0 && ny % 32 == 0) { *tx = 32; *ty = 32; }
    else if (Q == 27 && !means && nx % 32 == 0 && ny % 16 == 0) { *tx = 32; *ty = 16; }
    else if (Q == 27 && fp16 && nx % 32 == 0 && ny % 8 == 0) { *tx = 32; *ty = 8; }
    else if (nx % 16 == 0 && ny % 16 == 0) { *tx = 16; *ty = 16; }
    else if (nx % 32 == 0 && ny % 8 == 0) { *tx = 32; *ty = 8; }
    else { *tx = 8; *ty = 8; }
}

// Bytes per node of the FP64 running means a function-object mask turns on.
int mean_bytes(unsigned fo_mask) {
    return ((fo_mask & LBM_FO_KMEAN) ? 16 : 0) + ((fo_mask & LBM_FO_SMEAN) ? 96 : 0) +
           ((fo_mask & LBM_FO_MOMENTSMEAN) ? 160 : 0);
}

// Automatic tiling of a case with a function-object mask.
void auto_tiling(const lbm_config &c, unsigned fo_mask, int *tx, int *ty, int *lz) {
    const bool fp16 = c.storage == LBM_STORE_FP16;
    const int mb = mean_bytes(fo_mask);
    choose_tile(c.velocity_set, c.nx, c.ny, fp16, mb > 0, tx, ty);
    int64_t slots = resident_workgroups(c.velocity_set, *tx, *ty, c.device, c.storage, mb > 0);
    *lz = choose_lz(c.velocity_set, c.nx, c.ny, c.nz, *tx, *ty, slots, fp16, mb);
    // D3Q27 FP32 on 32x16 over more than one round of resident workgroups with a
    // partial last one: 32x32 (one workgroup per CU) measured faster (384^3:
    // 996 vs 1055 us); in whole rounds 32x16 stays ahead (256^3: 265 vs 272 us)
    // and below one round it has twice the work items to spread
    const int64_t items = (int64_t)(c.nx / *tx) * (c.ny / *ty) * (c.nz / *lz);
    if (c.velocity_set == 27 && !fp16 && mb == 0 && *tx == 32 && *ty == 16 && c.ny % 32 == 0 && items > slots &&
        items % slots != 0) {
        *ty = 32;
        slots = resident_workgroups(27, *tx, *ty, c.device, c.storage, false);
        *lz = choose_lz(27, c.nx, c.ny, c.nz, *tx, *ty, slots, false, 0);
    }
}

// Internal (column-tiled) index of a node.
static inline int64_t internal_index(int x, int y, int z, int nx, int ny, int tx, int ty) {
    const int nbx = nx / tx, nby = ny / ty;
    const int bx = x / tx, by = y / ty;
    return ((((int64_t)z * nby + by) * nbx + bx) * (tx * ty)) + (int64_t)(y % ty) * tx + (x % tx);
}

// perm[u] = internal index of the node at user-layout position u.
std::vector<int64_t> make_perm(int layout, int nx, int ny, int nz, int tx, int ty) {
    std::vector<int64_t> perm((size_t)nx * ny * nz);
    if (layout == LBM_LAYOUT_GLOBAL) {
        size_t u = 0;
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) perm[u++] = internal_index(x, y, z, nx, ny, tx, ty);
    } else {
        // reference 8^3 tiles: tx + 8(ty + 8(tz + 8(bx + NBX(by + NBY bz))))
        const int NBX = nx / 8, NBY = ny / 8, NBZ = nz / 8;
        size_t u = 0;
        for (int bz = 0; bz < NBZ; bz++)
            for (int by = 0; by < NBY; by++)
                for (int bx = 0; bx < NBX; bx++)
                    for (int tz = 0; tz < 8; tz++)
                        for (int ty8 = 0; ty8 < 8; ty8++)
                            for (int tx8 = 0; tx8 < 8; tx8++)
                                perm[u++] = internal_index(bx * 8 + tx8, by * 8 + ty8, bz * 8 + tz, nx, ny, tx, ty);
    }
    return perm;
}

}  // namespace lbm

// Function objects whose FP64 running means are fused into the step.
constexpr unsigned kMeanBits = LBM_FO_KMEAN | LBM_FO_SMEAN | LBM_FO_MOMENTSMEAN;

// Device allocation whose out-of-memory failure is reported as LBM_ERR_NOMEM.
static void dev_alloc(void **p, size_t bytes, const char *what) {
    const hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    *p = nullptr;
    if (e == hipErrorOutOfMemory)
        throw ArgError{LBM_ERR_NOMEM, std::string("device out of memory allocating ") + what};
    throw HipError{e, what};
}

struct lbm_handle {
    lbm_config cfg{};
    int Q = 19, TX = 16, TY = 16, LZ = 16, nbx = 0, nby = 0, nc = 0;
    bool auto_tile = false;  // tiling chosen by the library (may change with the function objects)
    int64_t n = 0;
    double omega = 0, tau = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void *mom[10] = {};    // the ten moment arrays, inside one allocation (mom_base)
    void *mom_base = nullptr;
    int mplane = 0;        // node words from a plane block of a moment to its next (StepParams::mplane)
    float *faces[2] = {nullptr, nullptr};
    int64_t fwords = 0, yoff = 0, zoff = 0;
    int parity = 0;
    unsigned fo_mask = 0;
    float *k = nullptr, *S = nullptr;
    double *mean_buf = nullptr;  // the enabled FP64 means, interleaved per plane block (StepParams)
    int mean_comp = 0;        // components in `means`
    StepClock *clock = nullptr;
    int64_t t = 0;
    float *uprev = nullptr;
    double *part = nullptr, *red_out = nullptr;
    int red_blocks = 1024;  // equal to kScratchBlocks: lbm_reduce_noalloc sums in the same order
    double *scratch = nullptr; // this device's copy of `scratch` (lbm_reduce_noalloc)
    bool red_first = true;
    std::map<int, hipGraphExec_t> graphs;  // key: parity*2 + means flag, for kGraphSteps steps
    std::map<int, std::vector<int64_t>> perms;
    std::string err;

    const std::vector<int64_t> &perm(int layout) {
        auto it = perms.find(layout);
        if (it == perms.end()) it = perms.emplace(layout, make_perm(layout, cfg.nx, cfg.ny, cfg.nz, TX, TY)).first;
        return it->second;
    }

    StepParams params(bool probe = false) const {
        StepParams p;
        std::memset(&p, 0, sizeof p);
        p.nx = cfg.nx;
        p.ny = cfg.ny;
        p.nz = cfg.nz;
        p.lz = LZ;
        p.nbx = nbx;
        p.nby = nby;
        p.nc = nc;
        p.omega = (float)omega;
        p.t_omega = (float)(1.0 - omega);
        p.omega_h = (float)(omega * 0.5);
        p.u_lid = (float)cfg.u_inf;
        p.tau = (float)tau;
        for (int i = 0; i < 10; i++) p.mom[i] = mom[i];
        p.mplane = mplane;
        p.mshift = __builtin_ctz((unsigned)(TX * TY));
        p.fin = faces[parity];
        p.fout = faces[parity ^ 1];
        p.yoff = yoff;
        p.zoff = zoff;
        codec(p.enc, p.dec);
        p.fo_mask = probe ? 0u : fo_mask;
        const int nt = TX * TY;
        int c = 0;
        p.kmean = (fo_mask & LBM_FO_KMEAN) ? mean_buf + (c++) * (size_t)nt : nullptr;
        p.smean = (fo_mask & LBM_FO_SMEAN) ? mean_buf + c * (size_t)nt : nullptr;
        c += (fo_mask & LBM_FO_SMEAN) ? 6 : 0;
        p.mmean = (fo_mask & LBM_FO_MOMENTSMEAN) ? mean_buf + c * (size_t)nt : nullptr;
        p.fcomp = nt;
        p.fplane = mean_comp * nt;
        p.clock = clock + parity;
        p.clock_next = clock + (parity ^ 1);
        p.s_diag = (float)(1.0 / (3.0 * tau));
        p.s_off = (float)(1.0 / (6.0 * tau));
        p.n = n;
        return p;
    }
    void codec(float *enc, float *dec) const {
        // FP16 storage: the step's compile-time codec (step_kernel.hpp)
        const bool h16 = cfg.storage == LBM_STORE_FP16;
        for (int i = 0; i < 10; i++) {
            enc[i] = h16 ? codec_enc(i) : 1.f;
            dec[i] = h16 ? codec_dec(i) : 1.f;
        }
    }
    int grid() const { return nbx * nby * nc; }
    // the step carries the FP64 time means (instantaneous k/S are formed on read)
    bool means() const { return (fo_mask & kMeanBits) != 0; }
    size_t elem() const { return cfg.storage == LBM_STORE_FP16 ? 2 : 4; }

    void drop_graphs() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    void release() {
        drop_graphs();
        if (mom_base) (void)hipFree(mom_base);
        mom_base = nullptr;
        for (auto &m : mom) m = nullptr;
        for (auto &f : faces) {
            if (f) (void)hipFree(f);
            f = nullptr;
        }
        for (void *p : {(void *)k, (void *)S, (void *)mean_buf, (void *)clock, (void *)uprev, (void *)part,
                        (void *)red_out})
            if (p) (void)hipFree(p);
        k = S = nullptr;
        mean_buf = nullptr;
        mean_comp = 0;
        clock = nullptr;
        uprev = nullptr;
        part = red_out = nullptr;
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

void lbm::handle_dims(const lbm_handle *h, int *nx, int *ny, int *nz) {
    *nx = h->cfg.nx;
    *ny = h->cfg.ny;
    *nz = h->cfg.nz;
}

namespace {

constexpr int kGraphSteps = 64;

int fail(lbm_handle *h, const HipError &e) {
    if (h) h->err = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
    return LBM_ERR_HIP;
}
int fail(lbm_handle *h, const ArgError &e) {
    if (h) h->err = e.msg;
    return e.code;
}

template <typename F>
int guarded(lbm_handle *h, F &&f) {
    if (!h) return LBM_ERR_ARG;
    try {
        HIPX(hipSetDevice(h->cfg.device));
        f();
        return LBM_OK;
    } catch (const HipError &e) {
        return fail(h, e);
    } catch (const ArgError &e) {
        return fail(h, e);
    } catch (const std::bad_alloc &) {
        h->err = "host allocation failed";
        return LBM_ERR_NOMEM;
    } catch (...) {
        h->err = "unexpected exception";
        return LBM_ERR_ARG;
    }
}

// One moment between a host array in the internal node order and its device
// array: plane blocks of tx ty elements, mplane elements apart on the device.
void copy_moment_h2d(lbm_handle *h, int k, const float *src) {
    const size_t nt = (size_t)h->TX * h->TY;
    HIPX(hipMemcpy2DAsync(h->mom[k], sizeof(float) * h->mplane, src, sizeof(float) * nt, sizeof(float) * nt,
                          (size_t)h->n / nt, hipMemcpyHostToDevice, h->stream));
}
void copy_moment_d2h(lbm_handle *h, int k, float *dst) {
    const size_t nt = (size_t)h->TX * h->TY;
    HIPX(hipMemcpy2DAsync(dst, sizeof(float) * nt, h->mom[k], sizeof(float) * h->mplane, sizeof(float) * nt,
                          (size_t)h->n / nt, hipMemcpyDeviceToHost, h->stream));
}

void upload_moment(lbm_handle *h, int k, const float *src, const std::vector<int64_t> &perm) {
    std::vector<float> tmp((size_t)h->n);
    for (int64_t u = 0; u < h->n; u++) tmp[(size_t)perm[(size_t)u]] = src[u];
    if (h->cfg.storage == LBM_STORE_FP32) {
        copy_moment_h2d(h, k, tmp.data());
        HIPX(hipStreamSynchronize(h->stream));
    } else {
        float *stage = nullptr;
        HIPX(hipMalloc(&stage, sizeof(float) * h->n));
        float enc[10], dec[10];
        h->codec(enc, dec);
        hipError_t e = hipMemcpyAsync(stage, tmp.data(), sizeof(float) * h->n, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(encode_f16_kernel, dim3(2048), dim3(256), 0, h->stream, h->params(), k, stage, enc[k]);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(stage);
        HIPX(e);
    }
}

void download_moment(lbm_handle *h, int k, float *dst, const std::vector<int64_t> &perm) {
    std::vector<float> tmp((size_t)h->n);
    if (h->cfg.storage == LBM_STORE_FP32) {
        copy_moment_d2h(h, k, tmp.data());
        HIPX(hipStreamSynchronize(h->stream));
    } else {
        float *stage = nullptr;
        HIPX(hipMalloc(&stage, sizeof(float) * h->n));
        float enc[10], dec[10];
        h->codec(enc, dec);
        hipLaunchKernelGGL(decode_f16_kernel, dim3(2048), dim3(256), 0, h->stream, h->params(), k, stage, dec[k]);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(tmp.data(), stage, sizeof(float) * h->n, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(stage);
        HIPX(e);
    }
    for (int64_t u = 0; u < h->n; u++) dst[u] = tmp[(size_t)perm[(size_t)u]];
}

void rebuild_faces(lbm_handle *h) {
    StepParams p = h->params();
    p.fout = h->faces[h->parity];  // the set the next step reads
    launch_fill(h->Q, h->TX, h->TY, p, h->grid(), h->stream, h->cfg.storage, false);
}

void check_layout(lbm_handle *h, int layout) {
    if (layout != LBM_LAYOUT_GLOBAL && layout != LBM_LAYOUT_TILED8) throw ArgError{LBM_ERR_ARG, "bad layout"};
    if (layout == LBM_LAYOUT_TILED8 && (h->cfg.nx % 8 || h->cfg.ny % 8 || h->cfg.nz % 8))
        throw ArgError{LBM_ERR_SIZE, "TILED8 layout needs nx, ny, nz multiples of 8"};
}

// Device step clock of the next step (the one of the current face parity) <-
// the handle's time t (stream-ordered; the caller syncs).
void set_clock(lbm_handle *h) {
    static thread_local StepClock c;  // stays valid until the copy has run
    c.t = h->t;
    c.inv = 1.0 / double(h->t + 1);
    HIPX(hipMemcpyAsync(h->clock + h->parity, &c, sizeof c, hipMemcpyHostToDevice, h->stream));
    HIPX(hipStreamSynchronize(h->stream));
}

// One step, launched on the handle's stream (captured or not).
void enqueue_step(lbm_handle *h, int parity) {
    StepParams p = h->params();
    p.fin = h->faces[parity];
    p.fout = h->faces[parity ^ 1];
    p.clock = h->clock + parity;
    p.clock_next = h->clock + (parity ^ 1);
    launch_step(h->Q, h->TX, h->TY, p, h->grid(), h->stream, h->cfg.storage, h->means(), false);
}

hipGraphExec_t graph_for(lbm_handle *h) {
    const int key = h->parity * 2 + (h->means() ? 1 : 0);
    auto it = h->graphs.find(key);
    if (it != h->graphs.end()) return it->second;
    hipGraph_t g = nullptr;
    HIPX(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int i = 0; i < kGraphSteps; i++) enqueue_step(h, h->parity ^ (i & 1));
    } catch (...) {
        hipGraph_t dummy = nullptr;
        (void)hipStreamEndCapture(h->stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
    }
    HIPX(hipStreamEndCapture(h->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPX(e);
    h->graphs[key] = ge;
    return ge;
}

void do_steps(lbm_handle *h, int n) {
    int left = n;
    while (left >= kGraphSteps) {
        hipGraphExec_t ge = graph_for(h);
        HIPX(hipGraphLaunch(ge, h->stream));
        left -= kGraphSteps;  // even count: parity unchanged
    }
    for (int i = 0; i < left; i++) {
        enqueue_step(h, h->parity);
        h->parity ^= 1;
    }
    h->t += n;
}

}  // namespace

void lbm::warm_moments(lbm_handle *h, const float *field, int S, int cstride, double U) {
    HIPX(hipSetDevice(h->cfg.device));
    HIPX(hipStreamSynchronize(h->stream));
    const StepParams p = h->params();
    const ColumnTiling t{h->TX, h->TY, h->nbx, h->nby, h->TX * h->TY};
    if (h->cfg.storage == LBM_STORE_FP32)
        hipLaunchKernelGGL(warm_moments_kernel<float>, dim3(4096), dim3(256), 0, h->stream, p, t, field, S, cstride, U);
    else
        hipLaunchKernelGGL(warm_moments_kernel<_Float16>, dim3(4096), dim3(256), 0, h->stream, p, t, field, S, cstride,
                           U);
    HIPX(hipGetLastError());
    rebuild_faces(h);
    HIPX(hipStreamSynchronize(h->stream));
}

extern "C" {

const char *lbm_create_error(void) { return g_create_error.c_str(); }

// Layout of the ten moment arrays of a tiling with tx ty nodes per plane
// block, in one allocation of `bytes`.  The arrays are made of 4-byte node
// words: one FP32 moment each (10 arrays), or with FP16 storage the codes of
// moments 2j and 2j+1 of a node in one word (5 arrays; moment 2j+1 two bytes
// after moment 2j).  Interleaved per plane block (mplane = arrays x tx ty node
// words) when every byte offset from mom[0] fits in 32 bits, else separate
// arrays back to back (mplane = tx ty, 256-byte aligned).  Interleaving makes
// each workgroup plane of all ten moments one contiguous run: a bare stream in
// the step's access shape measured 5.60 TB/s against 5.45 TB/s for ten arrays.
// kstride: bytes from one word array to the next.
static void moment_layout(const lbm_handle *h, int tx, int ty, int *mplane, size_t *kstride, size_t *bytes) {
    const size_t nt = (size_t)tx * ty, arrays = h->elem() == 4 ? 10 : 5, total = 4 * (size_t)h->n * arrays;
    if (total <= ((size_t)1 << 32)) {
        *mplane = (int)(arrays * nt);
        *kstride = 4 * nt;  // word array a = base + a tx ty words
        *bytes = total;
    } else {
        const size_t arr = (4 * (size_t)h->n + 255) / 256 * 256;
        *mplane = (int)nt;
        *kstride = arr;
        *bytes = arr * arrays;
    }
}
// Address of moment k in a layout of moment_layout.
static void *moment_ptr(const lbm_handle *h, void *base, size_t kstride, int k) {
    return h->elem() == 4 ? static_cast<char *>(base) + kstride * k
                          : static_cast<char *>(base) + kstride * (k / 2) + 2 * (k % 2);
}

// Allocates the moments (at rest) for the handle's current tiling.
static void alloc_moments(lbm_handle *h) {
    int mplane;
    size_t kstride, bytes;
    moment_layout(h, h->TX, h->TY, &mplane, &kstride, &bytes);
    void *base = nullptr;
    dev_alloc(&base, bytes, "the moments");
    const hipError_t e = hipMemsetAsync(base, 0, bytes, h->stream);
    if (e != hipSuccess) {
        (void)hipFree(base);
        HIPX(e);
    }
    h->mom_base = base;
    h->mplane = mplane;
    for (int i = 0; i < 10; i++) h->mom[i] = moment_ptr(h, base, kstride, i);
}

// Moves the case to another automatic tiling (footprint tx x ty, lz planes):
// the moments (and the residual's previous velocity field) are permuted on the
// device into the new column-tiled order, the face buffers are reallocated for
// the new geometry and rebuilt from the moments.  Everything new is allocated
// before anything old is freed; if that memory is not available the case keeps
// its tiling (a speed choice only) and false is returned.
static bool retile(lbm_handle *h, int tx, int ty, int lz, unsigned fo_mask) {
    const lbm_config &c = h->cfg;
    int64_t yoff = 0, zoff = 0;
    const int64_t fwords = face_words(h->Q, c.nx, c.ny, c.nz, tx, ty, lz, &yoff, &zoff);
    if (fwords >= (int64_t)INT32_MAX) return false;
    (void)fo_mask;
    int mplane;
    size_t kstride, bytes;
    moment_layout(h, tx, ty, &mplane, &kstride, &bytes);
    void *base = nullptr;
    float *faces[2] = {nullptr, nullptr}, *uprev = nullptr;
    auto free_new = [&] {
        for (void *p : {base, (void *)faces[0], (void *)faces[1], (void *)uprev})
            if (p) (void)hipFree(p);
    };
    auto get = [&](void **p, size_t bytes) {
        const hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) return true;
        (void)hipGetLastError();
        *p = nullptr;
        if (e != hipErrorOutOfMemory) {
            free_new();
            throw HipError{e, "hipMalloc (retile)"};
        }
        return false;
    };
    if (!get(&base, bytes) || !get((void **)&faces[0], sizeof(float) * fwords) ||
        !get((void **)&faces[1], sizeof(float) * fwords) ||
        (h->uprev && !get((void **)&uprev, sizeof(float) * 3 * h->n))) {
        free_new();
        return false;
    }
    const ColumnTiling from{h->TX, h->TY, h->nbx, h->nby, h->mplane}, to{tx, ty, c.nx / tx, c.ny / ty, mplane};
    const ColumnTiling from_plain{h->TX, h->TY, h->nbx, h->nby, h->TX * h->TY},
        to_plain{tx, ty, c.nx / tx, c.ny / ty, tx * ty};
    try {
        // node words: ten FP32 arrays, or five FP16 pair arrays (moments 2j, 2j+1)
        const int step = h->elem() == 4 ? 1 : 2;
        for (int k = 0; k < 10; k += step) {
            hipLaunchKernelGGL(retile_kernel<uint32_t>, dim3(4096), dim3(256), 0, h->stream,
                               (const uint32_t *)h->mom[k], (uint32_t *)moment_ptr(h, base, kstride, k), h->n, from,
                               to);
            HIPX(hipGetLastError());
        }
        for (int a = 0; uprev && a < 3; a++) {
            hipLaunchKernelGGL(retile_kernel<float>, dim3(4096), dim3(256), 0, h->stream, h->uprev + a * h->n,
                               uprev + a * h->n, h->n, from_plain, to_plain);
            HIPX(hipGetLastError());
        }
        for (int s = 0; s < 2; s++) HIPX(hipMemsetAsync(faces[s], 0, sizeof(float) * fwords, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
    } catch (...) {
        (void)hipStreamSynchronize(h->stream);
        free_new();
        throw;
    }
    (void)hipFree(h->mom_base);
    for (auto &f : h->faces) (void)hipFree(f);
    if (h->uprev) (void)hipFree(h->uprev);
    h->mom_base = base;
    h->mplane = mplane;
    for (int k = 0; k < 10; k++) h->mom[k] = moment_ptr(h, base, kstride, k);
    h->faces[0] = faces[0];
    h->faces[1] = faces[1];
    if (uprev) h->uprev = uprev;
    h->TX = tx;
    h->TY = ty;
    h->LZ = lz;
    h->nbx = to.nbx;
    h->nby = to.nby;
    h->nc = c.nz / lz;
    h->fwords = fwords;
    h->yoff = yoff;
    h->zoff = zoff;
    h->perms.clear();
    h->drop_graphs();
    rebuild_faces(h);
    HIPX(hipStreamSynchronize(h->stream));
    return true;
}

int lbm_create(const lbm_config *cfg, lbm_handle **out) {
    g_create_error.clear();
    if (!cfg || !out) {
        g_create_error = "null argument";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    lbm_handle *h = new (std::nothrow) lbm_handle();
    if (!h) return LBM_ERR_NOMEM;
    h->cfg = *cfg;
    const lbm_config &c = h->cfg;
    auto bad = [&](int code, const std::string &msg) {
        g_create_error = msg;
        h->release();
        delete h;
        return code;
    };
    if (c.velocity_set != 19 && c.velocity_set != 27) return bad(LBM_ERR_ARG, "velocity_set must be 19 or 27");
    if (c.storage != LBM_STORE_FP32 && c.storage != LBM_STORE_FP16) return bad(LBM_ERR_ARG, "bad storage");
    if (c.nx < 2 || c.ny < 2 || c.nz < 2) return bad(LBM_ERR_SIZE, "mesh dimensions must be >= 2");
    if (!(c.Re > 0) || !std::isfinite(c.u_inf)) return bad(LBM_ERR_ARG, "Re must be > 0 and u_inf finite");
    h->Q = c.velocity_set;
    h->auto_tile = (c.tile_x == 0 || c.tile_y == 0) && c.tile_z <= 0;
    if (c.tile_x == 0 || c.tile_y == 0) {
        int lz;
        auto_tiling(c, 0u, &h->TX, &h->TY, &lz);
    } else {
        h->TX = c.tile_x;
        h->TY = c.tile_y;
    }
    if (!tile_supported(h->Q, h->TX, h->TY)) return bad(LBM_ERR_ARG, "unsupported tile footprint");
    if (c.nx % h->TX || c.ny % h->TY)
        return bad(LBM_ERR_SIZE, "nx and ny must be multiples of the tile footprint " + std::to_string(h->TX) +
                                     "x" + std::to_string(h->TY));
    h->n = (int64_t)c.nx * c.ny * c.nz;
    // moment arrays are addressed with 32-bit byte offsets (step_kernel.hpp, Store)
    if (h->n > ((int64_t)1 << 30)) return bad(LBM_ERR_SIZE, "mesh too large: at most 2^30 nodes (4 GiB per moment array)");
    h->LZ = c.tile_z > 0 ? c.tile_z
                         : choose_lz(h->Q, c.nx, c.ny, c.nz, h->TX, h->TY,
                                     resident_workgroups(h->Q, h->TX, h->TY, c.device, c.storage, false),
                                     c.storage == LBM_STORE_FP16, 0);
    if (c.nz % h->LZ) return bad(LBM_ERR_SIZE, "nz must be a multiple of tile_z");
    h->nbx = c.nx / h->TX;
    h->nby = c.ny / h->TY;
    h->nc = c.nz / h->LZ;
    relaxation(c.Re, c.u_inf, c.nx, &h->omega, &h->tau);
    try {
        HIPX(hipSetDevice(c.device));
        if (c.stream)
            h->stream = (hipStream_t)c.stream;
        else {
            HIPX(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            h->own_stream = true;
        }
        alloc_moments(h);
        h->fwords = face_words(h->Q, c.nx, c.ny, c.nz, h->TX, h->TY, h->LZ, &h->yoff, &h->zoff);
        if (h->fwords >= (int64_t)INT32_MAX) throw ArgError{LBM_ERR_SIZE, "face buffers exceed 2^31 words"};
        for (int s = 0; s < 2; s++) {
            dev_alloc((void **)&h->faces[s], sizeof(float) * h->fwords, "face buffers");
            HIPX(hipMemsetAsync(h->faces[s], 0, sizeof(float) * h->fwords, h->stream));
        }
        dev_alloc((void **)&h->clock, 2 * sizeof(StepClock), "step clock");
        h->scratch = scratch_on(c.device);
        set_clock(h);
        rebuild_faces(h);
        HIPX(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return bad(LBM_ERR_HIP, std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what);
    } catch (const ArgError &e) {
        return bad(e.code, e.msg);
    }
    *out = h;
    return LBM_OK;
}

int lbm_destroy(lbm_handle *h) {
    if (!h) return LBM_ERR_ARG;
    (void)hipSetDevice(h->cfg.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->release();
    delete h;
    return LBM_OK;
}

const char *lbm_last_error(const lbm_handle *h) { return h ? h->err.c_str() : "null handle"; }

int lbm_get_tiling(const lbm_handle *h, int *tx, int *ty, int *tz) {
    if (!h) return LBM_ERR_ARG;
    if (tx) *tx = h->TX;
    if (ty) *ty = h->TY;
    if (tz) *tz = h->LZ;
    return LBM_OK;
}

int lbm_get_omega(const lbm_handle *h, double *omega, double *tau) {
    if (!h) return LBM_ERR_ARG;
    if (omega) *omega = h->omega;
    if (tau) *tau = h->tau;
    return LBM_OK;
}

int lbm_set_moments(lbm_handle *h, const float *const m[10], int layout) {
    return guarded(h, [&] {
        if (!m) throw ArgError{LBM_ERR_ARG, "null moment array"};
        check_layout(h, layout);
        HIPX(hipStreamSynchronize(h->stream));
        const auto &perm = h->perm(layout);
        for (int k = 0; k < 10; k++) {
            if (!m[k]) throw ArgError{LBM_ERR_ARG, "null moment array"};
            upload_moment(h, k, m[k], perm);
        }
        rebuild_faces(h);
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int lbm_get_moments(lbm_handle *h, float *m[10], int layout) {
    return guarded(h, [&] {
        if (!m) throw ArgError{LBM_ERR_ARG, "null moment array"};
        check_layout(h, layout);
        HIPX(hipStreamSynchronize(h->stream));
        const auto &perm = h->perm(layout);
        for (int k = 0; k < 10; k++) {
            if (!m[k]) throw ArgError{LBM_ERR_ARG, "null moment array"};
            download_moment(h, k, m[k], perm);
        }
    });
}

int lbm_set_moments_f64(lbm_handle *h, const double *const m[10], int layout) {
    if (!h || !m) return LBM_ERR_ARG;
    std::vector<std::vector<float>> tmp(10);
    const float *ptrs[10];
    for (int k = 0; k < 10; k++) {
        if (!m[k]) return LBM_ERR_ARG;
        tmp[k].resize((size_t)h->n);
        for (int64_t i = 0; i < h->n; i++) tmp[k][(size_t)i] = (float)m[k][i];
        ptrs[k] = tmp[k].data();
    }
    return lbm_set_moments(h, ptrs, layout);
}

int lbm_get_moments_f64(lbm_handle *h, double *m[10], int layout) {
    if (!h || !m) return LBM_ERR_ARG;
    std::vector<std::vector<float>> tmp(10);
    float *ptrs[10];
    for (int k = 0; k < 10; k++) {
        tmp[k].resize((size_t)h->n);
        ptrs[k] = tmp[k].data();
    }
    const int rc = lbm_get_moments(h, ptrs, layout);
    if (rc) return rc;
    for (int k = 0; k < 10; k++) {
        if (!m[k]) return LBM_ERR_ARG;
        for (int64_t i = 0; i < h->n; i++) m[k][i] = tmp[k][(size_t)i];
    }
    return LBM_OK;
}

int lbm_step(lbm_handle *h, int n) {
    return guarded(h, [&] {
        if (n < 0) throw ArgError{LBM_ERR_ARG, "negative step count"};
        do_steps(h, n);
    });
}

int lbm_synchronize(lbm_handle *h) {
    return guarded(h, [&] { HIPX(hipStreamSynchronize(h->stream)); });
}

int lbm_step_timed(lbm_handle *h, int n, double *ms) {
    return guarded(h, [&] {
        if (n < 0 || !ms) throw ArgError{LBM_ERR_ARG, "bad argument"};
        // build graphs outside the timed region
        if (n >= kGraphSteps) {
            (void)graph_for(h);
        }
        hipEvent_t a, b;
        HIPX(hipEventCreate(&a));
        HIPX(hipEventCreate(&b));
        hipError_t e = hipEventRecord(a, h->stream);
        try {
            if (e == hipSuccess) do_steps(h, n);
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        if (e == hipSuccess) e = hipEventRecord(b, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        HIPX(e);
        *ms = f;
    });
}

int lbm_get_time(const lbm_handle *h, int64_t *t) {
    if (!h || !t) return LBM_ERR_ARG;
    *t = h->t;
    return LBM_OK;
}

int lbm_set_time(lbm_handle *h, int64_t t) {
    return guarded(h, [&] {
        if (t < 0) throw ArgError{LBM_ERR_ARG, "negative time"};
        h->t = t;
        set_clock(h);
        HIPX(hipStreamSynchronize(h->stream));
    });
}

int lbm_fo_enable(lbm_handle *h, unsigned mask) {
    return guarded(h, [&] {
        if (mask & ~31u) throw ArgError{LBM_ERR_ARG, "unknown function-object bits"};
        HIPX(hipStreamSynchronize(h->stream));
        h->drop_graphs();
        // an automatic tiling follows the function objects (with an explicit one,
        // or when the memory for the move is not available, nothing moves)
        if (h->auto_tile) {
            int tx, ty, lz;
            auto_tiling(h->cfg, mask, &tx, &ty, &lz);
            if (tx != h->TX || ty != h->TY || lz != h->LZ) (void)retile(h, tx, ty, lz, mask);
        }
        const int64_t n = h->n;
        if ((mask & LBM_FO_K) && !h->k) dev_alloc((void **)&h->k, sizeof(float) * n, "k");
        if ((mask & LBM_FO_S) && !h->S) dev_alloc((void **)&h->S, sizeof(float) * 6 * n, "S");
        // the means' components, interleaved per plane block; reset to zero
        const int comp = ((mask & LBM_FO_KMEAN) ? 1 : 0) + ((mask & LBM_FO_SMEAN) ? 6 : 0) +
                         ((mask & LBM_FO_MOMENTSMEAN) ? 10 : 0);
        if (comp != h->mean_comp) {
            if (h->mean_buf) (void)hipFree(h->mean_buf);
            h->mean_buf = nullptr;
            h->mean_comp = 0;
            if (comp) dev_alloc((void **)&h->mean_buf, sizeof(double) * comp * n, "the FP64 means");
            h->mean_comp = comp;
        }
        if (comp) {
            hipLaunchKernelGGL(fill_kernel_f64, dim3(2048), dim3(256), 0, h->stream, h->mean_buf, (int64_t)comp * n, 0.0);
            HIPX(hipGetLastError());
        }
        set_clock(h);
        HIPX(hipStreamSynchronize(h->stream));
        h->fo_mask = mask;
    });
}

int lbm_fo_get(lbm_handle *h, int field, double *out, int layout) {
    return guarded(h, [&] {
        if (!out) throw ArgError{LBM_ERR_ARG, "null output"};
        check_layout(h, layout);
        const unsigned need = field == LBM_FIELD_K ? LBM_FO_K
                              : field == LBM_FIELD_KMEAN ? LBM_FO_KMEAN
                              : field == LBM_FIELD_S ? LBM_FO_S
                              : field == LBM_FIELD_SMEAN ? LBM_FO_SMEAN
                              : field == LBM_FIELD_MOMENTSMEAN ? LBM_FO_MOMENTSMEAN : 0u;
        if (!need) throw ArgError{LBM_ERR_ARG, "unknown field"};
        if (!(h->fo_mask & need)) throw ArgError{LBM_ERR_STATE, "function object not enabled"};
        HIPX(hipStreamSynchronize(h->stream));
        const int ncomp = (field == LBM_FIELD_S || field == LBM_FIELD_SMEAN) ? 6
                          : field == LBM_FIELD_MOMENTSMEAN ? 10 : 1;
        const int64_t n = h->n;
        const auto &perm = h->perm(layout);
        std::vector<double> tmp((size_t)(n * ncomp));
        if (field == LBM_FIELD_K || field == LBM_FIELD_S) {
            // instantaneous fields from the current stored moments
            float *dk = field == LBM_FIELD_K ? h->k : nullptr, *dS = field == LBM_FIELD_S ? h->S : nullptr;
            const StepParams p = h->params();
            if (h->cfg.storage == LBM_STORE_FP32)
                hipLaunchKernelGGL(fo_instant_kernel<float>, dim3(2048), dim3(256), 0, h->stream, p, dk, dS);
            else
                hipLaunchKernelGGL(fo_instant_kernel<_Float16>, dim3(2048), dim3(256), 0, h->stream, p, dk, dS);
            HIPX(hipGetLastError());
            HIPX(hipStreamSynchronize(h->stream));
            std::vector<float> f((size_t)(n * ncomp));
            HIPX(hipMemcpy(f.data(), field == LBM_FIELD_K ? h->k : h->S, sizeof(float) * n * ncomp,
                           hipMemcpyDeviceToHost));
            for (size_t i = 0; i < f.size(); i++) tmp[i] = f[i];
        } else {
            // de-interleave the field's components from the plane blocks
            std::vector<double> all((size_t)n * h->mean_comp);
            HIPX(hipMemcpy(all.data(), h->mean_buf, sizeof(double) * all.size(), hipMemcpyDeviceToHost));
            const StepParams p = h->params();
            const double *first = field == LBM_FIELD_KMEAN ? p.kmean : field == LBM_FIELD_SMEAN ? p.smean : p.mmean;
            const int64_t c0 = (first - h->mean_buf) / p.fcomp, nt = p.fcomp;
            for (int c = 0; c < ncomp; c++)
                for (int64_t i = 0; i < n; i++)
                    tmp[(size_t)(c * n + i)] = all[(size_t)((i / nt) * p.fplane + (c0 + c) * nt + i % nt)];
        }
        for (int c = 0; c < ncomp; c++)
            for (int64_t u = 0; u < n; u++) out[c * n + u] = tmp[(size_t)(c * n + perm[(size_t)u])];
    });
}

int lbm_reduce(lbm_handle *h, double *ke, double *res) {
    return guarded(h, [&] {
        if (!h->uprev || !h->part || !h->red_out) {  // first call, or a first call that failed part-way
            if (!h->uprev) dev_alloc((void **)&h->uprev, sizeof(float) * 3 * h->n, "previous velocity");
            if (!h->part) dev_alloc((void **)&h->part, sizeof(double) * 3 * h->red_blocks, "reduction partials");
            if (!h->red_out) dev_alloc((void **)&h->red_out, sizeof(double) * 3, "reduction result");
            h->red_first = true;
        }
        StepParams p = h->params();
        constexpr int NT = 256;
        if (h->cfg.storage == LBM_STORE_FP32)
            hipLaunchKernelGGL((residual_partials_kernel<float, NT>), dim3(h->red_blocks), dim3(NT), 0, h->stream, p,
                               h->uprev, h->part, h->red_first ? 1 : 0);
        else
            hipLaunchKernelGGL((residual_partials_kernel<_Float16, NT>), dim3(h->red_blocks), dim3(NT), 0,
                               h->stream, p, h->uprev, h->part, h->red_first ? 1 : 0);
        HIPX(hipGetLastError());
        hipLaunchKernelGGL((residual_final_kernel<NT>), dim3(1), dim3(NT), 0, h->stream, h->part, h->red_blocks,
                           h->red_out);
        HIPX(hipGetLastError());
        double r[3];
        HIPX(hipMemcpyAsync(r, h->red_out, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIPX(hipStreamSynchronize(h->stream));
        h->red_first = false;
        if (ke) *ke = r[0];
        if (res) *res = r[2] > 0 ? std::sqrt(r[1] / r[2]) : (r[1] > 0 ? INFINITY : 0.0);
    });
}

int lbm_reduce_noalloc(lbm_handle *h, float *uprev, int64_t uprev_count, int first, double *out) {
    return guarded(h, [&] {
        if (!uprev || !out) throw ArgError{LBM_ERR_ARG, "null device buffer"};
        if (uprev_count < 3 * h->n)
            throw ArgError{LBM_ERR_SIZE, "previous-velocity buffer holds " + std::to_string(uprev_count) +
                                             " floats; this mesh needs 3 * nodes = " + std::to_string(3 * h->n)};
        if (h->red_blocks > kScratchBlocks)
            throw ArgError{LBM_ERR_SIZE, "reduction needs " + std::to_string(h->red_blocks) +
                                             " partial blocks; the static scratch holds " +
                                             std::to_string(kScratchBlocks)};
        if (!h->scratch) throw ArgError{LBM_ERR_STATE, "static scratch not available on this device"};
        double *part = h->scratch, *sums = h->scratch + 3 * kScratchBlocks;
        const StepParams p = h->params();
        constexpr int NT = 256;
        if (h->cfg.storage == LBM_STORE_FP32)
            hipLaunchKernelGGL((residual_partials_kernel<float, NT>), dim3(h->red_blocks), dim3(NT), 0, h->stream, p,
                               uprev, part, first ? 1 : 0);
        else
            hipLaunchKernelGGL((residual_partials_kernel<_Float16, NT>), dim3(h->red_blocks), dim3(NT), 0,
                               h->stream, p, uprev, part, first ? 1 : 0);
        HIPX(hipGetLastError());
        hipLaunchKernelGGL((residual_final_kernel<NT>), dim3(1), dim3(NT), 0, h->stream, part, h->red_blocks, sums);
        HIPX(hipGetLastError());
        hipLaunchKernelGGL(residual_result_kernel, dim3(1), dim3(64), 0, h->stream, sums, out);
        HIPX(hipGetLastError());
    });
}

int lbm_probe_streaming(lbm_handle *h, uint32_t *src_out, uint8_t *type_out) {
    return guarded(h, [&] {
        if (!src_out || !type_out) throw ArgError{LBM_ERR_ARG, "null output"};
        if (h->n * 32 + 31 >= (int64_t)UINT32_MAX) throw ArgError{LBM_ERR_SIZE, "mesh too large for 32-bit probe ids"};
        HIPX(hipStreamSynchronize(h->stream));
        const int64_t n = h->n;
        uint32_t *dsrc = nullptr;
        uint8_t *dtype = nullptr;
        HIPX(hipMalloc(&dsrc, sizeof(uint32_t) * n * h->Q));
        hipError_t e = hipMalloc(&dtype, n);
        if (e != hipSuccess) {
            (void)hipFree(dsrc);
            HIPX(e);
        }
        try {
            StepParams p = h->params(true);
            p.probe_src = dsrc;
            p.probe_type = dtype;
            // ids into the set the probe step reads, then one probe step
            StepParams pf = p;
            pf.fout = h->faces[h->parity];
            launch_fill(h->Q, h->TX, h->TY, pf, h->grid(), h->stream, LBM_STORE_FP32, true);
            launch_step(h->Q, h->TX, h->TY, p, h->grid(), h->stream, LBM_STORE_FP32, false, true);
            HIPX(hipMemcpyAsync(src_out, dsrc, sizeof(uint32_t) * n * h->Q, hipMemcpyDeviceToHost, h->stream));
            HIPX(hipMemcpyAsync(type_out, dtype, n, hipMemcpyDeviceToHost, h->stream));
            HIPX(hipStreamSynchronize(h->stream));
            // restore the face set from the moments
            rebuild_faces(h);
            HIPX(hipStreamSynchronize(h->stream));
        } catch (...) {
            (void)hipFree(dsrc);
            (void)hipFree(dtype);
            throw;
        }
        (void)hipFree(dsrc);
        (void)hipFree(dtype);
    });
}

}  // extern "C"
