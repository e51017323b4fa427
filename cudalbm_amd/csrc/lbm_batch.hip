// lbm_batch.hip -- C ABI of the batched multi-case sweep (lbm_batch_*) and of
// the single-case tile autotuner (lbm_autotune).
//
// A batch holds independent cavities of uneven size, Re and lid speed.  Every
// case keeps its own moments and two face-buffer sets, laid out exactly as a
// single-case solver's.  One step of the whole batch is ONE launch of
// batch_step_kernel over a tile-to-case table of (case, column, chunk) work
// items; the per-case parameters (omega, lid speed, buffer addresses) come from
// a device-side table with one row per (face parity, case).  Steps are replayed
// from hipGraphs.  Converged cases are dropped from the item table, so later
// launches cover only the cases still running.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "lbm_amd.h"
#include "aux_kernels.hpp"
#include "lbm_internal.hpp"

using namespace lbm;

namespace {

constexpr int kBatchGraphSteps = 64;
constexpr int kResBlocksPerCase = 64;  // FP64 residual partials per case
constexpr int kResThreads = 256;

// ------------------------------------------------------------ kernels
// Per-case partial sums: block b works on case b / kResBlocksPerCase.
template <int NT>
__global__ __launch_bounds__(NT) void batch_residual_partials_kernel(const StepParams *__restrict__ cases,
                                                                     float *const *__restrict__ uprev,
                                                                     const int *__restrict__ first,
                                                                     double *__restrict__ part) {
    __shared__ double sh[NT / 64];
    const int cs = blockIdx.x / kResBlocksPerCase, j = blockIdx.x % kResBlocksPerCase;
    residual_partials<float, NT>(cases[cs], uprev[cs], part + 3 * blockIdx.x, first[cs], j, kResBlocksPerCase, sh);
}

// ------------------------------------------------------------ dispatch
template <int Q, int TX, int TY, bool PERSIST>
void launch_batch_t(const StepParams *cases, const BatchItem *items, int count, int grid, hipStream_t s) {
    hipLaunchKernelGGL((batch_step_kernel<Q, TX, TY, PERSIST>), dim3(grid), dim3(TX * TY), 0, s, cases, items,
                       count);
}

template <int Q, int TX, int TY, bool PERSIST>
int occupancy_t() {
    int nb = 0;
    HIPX(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_step_kernel<Q, TX, TY, PERSIST>, TX * TY, 0));
    return nb;
}

// Calls F<Q,TX,TY,PERSIST>::run for a runtime (Q, tx, ty, persist); false when
// that combination is not built.  D3Q27 is built for 16x16 and for 8x8 without
// the persistent form (its item loop does not fit the register file without
// spilling).
template <template <int, int, int, bool> class F, typename... A>
bool dispatch(int Q, int tx, int ty, bool persist, A &&...a) {
#define LBM_TILE(QQ, X, Y)                                                   \
    if (Q == QQ && tx == X && ty == Y) {                                     \
        if (persist) F<QQ, X, Y, true>::run(a...);                           \
        else F<QQ, X, Y, false>::run(a...);                                  \
        return true;                                                         \
    }
    LBM_TILE(19, 16, 16)
    LBM_TILE(19, 32, 16)
    LBM_TILE(19, 32, 8)
    LBM_TILE(19, 8, 8)
    LBM_TILE(27, 16, 16)
#undef LBM_TILE
    if (Q == 27 && tx == 8 && ty == 8 && !persist) {
        F<27, 8, 8, false>::run(a...);
        return true;
    }
    return false;
}

template <int Q, int TX, int TY, bool PERSIST>
struct LaunchF {
    static void run(const StepParams *cases, const BatchItem *items, int count, int grid, hipStream_t s) {
        launch_batch_t<Q, TX, TY, PERSIST>(cases, items, count, grid, s);
    }
};
template <int Q, int TX, int TY, bool PERSIST>
struct OccF {
    static void run(int *nb) { *nb = occupancy_t<Q, TX, TY, PERSIST>(); }
};

bool batch_supported(int Q, int tx, int ty, bool persist) {
    if (Q == 19) return (tx == 16 && ty == 16) || (tx == 32 && ty == 8) || (tx == 8 && ty == 8) || (tx == 32 && ty == 16);
    return (tx == 16 && ty == 16) || (tx == 8 && ty == 8 && !persist);
}

// Largest divisor of nz not above target (at least 1).
int lz_for(int nz, int target) {
    for (int lz = std::min(nz, target); lz > 1; lz--)
        if (nz % lz == 0) return lz;
    return 1;
}

struct Case {
    lbm_batch_case c{};
    int lz = 1, nbx = 0, nby = 0, nc = 0;
    int64_t n = 0, fwords = 0, yoff = 0, zoff = 0;
    double omega = 0, tau = 0;
    float *mom = nullptr;    // the 10 moments of n nodes: interleaved per plane block (mplane())
    float *faces = nullptr;  // 2 face sets of fwords floats
    float *uprev = nullptr;  // 3 arrays of n floats (residual reference field)
    int64_t t = 0;
    bool active = true, red_first = true;

    // Parameters of a step that reads face set `set` (and writes the other).
    StepParams params(int set) const {
        StepParams p;
        std::memset(&p, 0, sizeof p);
        p.nx = c.nx;
        p.ny = c.ny;
        p.nz = c.nz;
        p.lz = lz;
        p.nbx = nbx;
        p.nby = nby;
        p.nc = nc;
        p.omega = (float)omega;
        p.t_omega = (float)(1.0 - omega);
        p.omega_h = (float)(omega * 0.5);
        p.u_lid = (float)c.u_inf;
        p.tau = (float)tau;
        // interleaved per plane block as a single case's moments (StepParams::mplane),
        // or ten separate arrays where 32-bit byte offsets would not reach
        for (int k = 0; k < 10; k++) p.mom[k] = mom + (int64_t)k * kstride();
        p.mplane = mplane();
        p.mshift = __builtin_ctz((unsigned)nt());
        p.fin = faces + (int64_t)set * fwords;
        p.fout = faces + (int64_t)(set ^ 1) * fwords;
        p.yoff = (int)yoff;
        p.zoff = (int)zoff;
        for (int k = 0; k < 10; k++) p.enc[k] = p.dec[k] = 1.f;
        p.n = n;
        return p;
    }
    int grid() const { return nbx * nby * nc; }
    int nt() const { return (c.nx / nbx) * (c.ny / nby); }
    bool interleaved() const { return sizeof(float) * 10 * (size_t)n <= ((size_t)1 << 32); }
    int mplane() const { return interleaved() ? 10 * nt() : nt(); }
    int64_t kstride() const { return interleaved() ? nt() : n; }  // elements from mom[k] to mom[k+1]
    // one moment between a host array in the internal node order and the device
    void h2d(int k, const float *src, hipStream_t s) const {
        const size_t w = sizeof(float) * nt();
        HIPX(hipMemcpy2DAsync(mom + (int64_t)k * kstride(), sizeof(float) * mplane(), src, w, w, (size_t)n / nt(),
                              hipMemcpyHostToDevice, s));
    }
    void d2h(int k, float *dst, hipStream_t s) const {
        const size_t w = sizeof(float) * nt();
        HIPX(hipMemcpy2DAsync(dst, w, mom + (int64_t)k * kstride(), sizeof(float) * mplane(), w, (size_t)n / nt(),
                              hipMemcpyDeviceToHost, s));
    }
};

}  // namespace

struct lbm_batch {
    lbm_batch_config cfg{};
    int Q = 19, TX = 16, TY = 16, LZT = 16;
    bool persist = false;
    int grid_persist = 0;  // workgroups of the persistent grid
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<Case> cases;
    // device tables: params[2][ncases] (row g: steps taken at global parity g),
    // the work items of the active cases, and the residual helpers
    StepParams *dtab = nullptr;
    BatchItem *ditems = nullptr;
    float **duprev = nullptr;
    int *dfirst = nullptr;
    double *dpart = nullptr, *dres = nullptr;
    int nitems = 0;
    bool tables_dirty = true;
    int gpar = 0;  // global parity: flips with every batch step
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (steps, gpar)
    std::string err, report;

    int ncases() const { return (int)cases.size(); }

    void drop_graphs() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    void release() {
        drop_graphs();
        for (auto &c : cases) {
            for (void *p : {(void *)c.mom, (void *)c.faces, (void *)c.uprev})
                if (p) (void)hipFree(p);
            c.mom = c.faces = c.uprev = nullptr;
        }
        for (void *p : {(void *)dtab, (void *)ditems, (void *)duprev, (void *)dfirst, (void *)dpart, (void *)dres})
            if (p) (void)hipFree(p);
        dtab = nullptr;
        ditems = nullptr;
        duprev = nullptr;
        dfirst = nullptr;
        dpart = dres = nullptr;
        if (own_stream && stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

thread_local std::string g_batch_create_error;

template <typename F>
int guarded(lbm_batch *b, F &&f) {
    if (!b) return LBM_ERR_ARG;
    try {
        HIPX(hipSetDevice(b->cfg.device));
        f();
        return LBM_OK;
    } catch (const HipError &e) {
        b->err = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
        return LBM_ERR_HIP;
    } catch (const ArgError &e) {
        b->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        b->err = "host allocation failed";
        return LBM_ERR_NOMEM;
    } catch (...) {
        b->err = "unexpected exception";
        return LBM_ERR_ARG;
    }
}

// Face set case i reads at its next step.
int next_set(const Case &c) { return (int)(c.t & 1); }

// Rewrites the parameter rows and the work-item table from the cases' state.
// Stream-ordered and synchronous; never called during capture.
void sync_tables(lbm_batch *b) {
    if (!b->tables_dirty) return;
    const int nc = b->ncases();
    std::vector<StepParams> tab(2 * (size_t)nc);
    for (int i = 0; i < nc; i++) {
        const Case &c = b->cases[i];
        // at global parity g the case reads set next_set ^ (g ^ gpar)
        for (int g = 0; g < 2; g++) tab[(size_t)g * nc + i] = c.params(next_set(c) ^ g ^ b->gpar);
    }
    // items of the active cases: deepest chunks first (they take longest), then
    // by case, then chunk-major inside a case
    std::vector<int> order;
    for (int i = 0; i < nc; i++)
        if (b->cases[i].active) order.push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return b->cases[a].lz > b->cases[c].lz; });
    std::vector<BatchItem> items;
    for (int i : order) {
        const Case &c = b->cases[i];
        for (int w = 0; w < c.grid(); w++) items.push_back({i, w % c.nbx, (w / c.nbx) % c.nby, w / (c.nbx * c.nby)});
    }
    HIPX(hipMemcpyAsync(b->dtab, tab.data(), sizeof(StepParams) * tab.size(), hipMemcpyHostToDevice, b->stream));
    if (!items.empty())
        HIPX(hipMemcpyAsync(b->ditems, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice,
                            b->stream));
    HIPX(hipStreamSynchronize(b->stream));
    if ((int)items.size() != b->nitems) b->drop_graphs();  // grid / count are part of the graphs
    b->nitems = (int)items.size();
    b->tables_dirty = false;
}

void enqueue_step(lbm_batch *b, int g) {
    if (b->nitems == 0) return;
    const int grid = b->persist ? std::min(b->grid_persist, b->nitems) : b->nitems;
    if (!dispatch<LaunchF>(b->Q, b->TX, b->TY, b->persist, b->dtab + (size_t)g * b->ncases(), b->ditems,
                           b->nitems, grid, b->stream))
        throw ArgError{LBM_ERR_ARG, "unsupported batch tile"};
    HIPX(hipGetLastError());
}

hipGraphExec_t graph_for(lbm_batch *b, int steps) {
    const auto key = std::make_pair(steps, b->gpar);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) return it->second;
    hipGraph_t g = nullptr;
    HIPX(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
        for (int i = 0; i < steps; i++) enqueue_step(b, b->gpar ^ (i & 1));
    } catch (...) {
        hipGraph_t dummy = nullptr;
        (void)hipStreamEndCapture(b->stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
    }
    HIPX(hipStreamEndCapture(b->stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPX(e);
    b->graphs[key] = ge;
    return ge;
}

void prepare_graphs(lbm_batch *b, int n) {
    sync_tables(b);
    if (b->nitems == 0) return;
    if (n >= kBatchGraphSteps) (void)graph_for(b, kBatchGraphSteps);
    // full graphs have an even step count, so the remainder starts at gpar too
    if (n % kBatchGraphSteps) (void)graph_for(b, n % kBatchGraphSteps);
}

void do_steps(lbm_batch *b, int n) {
    prepare_graphs(b, n);
    if (b->nitems == 0 || n == 0) return;
    int left = n;
    while (left > 0) {
        const int k = std::min(left, kBatchGraphSteps);
        HIPX(hipGraphLaunch(graph_for(b, k), b->stream));
        b->gpar ^= k & 1;
        left -= k;
    }
    for (auto &c : b->cases)
        if (c.active) c.t += n;
}

void fill_faces(lbm_batch *b, int i) {
    const Case &c = b->cases[i];
    StepParams p = c.params(next_set(c) ^ 1);  // fout = the set the next step reads
    launch_fill(b->Q, b->TX, b->TY, p, c.grid(), b->stream, LBM_STORE_FP32, false);
}

void check_case(lbm_batch *b, int cs) {
    if (cs < 0 || cs >= b->ncases()) throw ArgError{LBM_ERR_ARG, "case index out of range"};
}

void check_layout(const Case &c, int layout) {
    if (layout != LBM_LAYOUT_GLOBAL && layout != LBM_LAYOUT_TILED8) throw ArgError{LBM_ERR_ARG, "bad layout"};
    if (layout == LBM_LAYOUT_TILED8 && (c.c.nx % 8 || c.c.ny % 8 || c.c.nz % 8))
        throw ArgError{LBM_ERR_SIZE, "TILED8 layout needs nx, ny, nz multiples of 8"};
}

void reduce_all(lbm_batch *b, std::vector<double> &ke, std::vector<double> &res) {
    const int nc = b->ncases();
    if (!b->duprev) {
        std::vector<float *> ptrs(nc);
        for (int i = 0; i < nc; i++) {
            Case &c = b->cases[i];
            HIPX(hipMalloc(&c.uprev, sizeof(float) * 3 * c.n));
            ptrs[i] = c.uprev;
        }
        HIPX(hipMalloc(&b->duprev, sizeof(float *) * nc));
        HIPX(hipMalloc(&b->dfirst, sizeof(int) * nc));
        HIPX(hipMalloc(&b->dpart, sizeof(double) * 3 * kResBlocksPerCase * nc));
        HIPX(hipMalloc(&b->dres, sizeof(double) * 3 * nc));
        HIPX(hipMemcpyAsync(b->duprev, ptrs.data(), sizeof(float *) * nc, hipMemcpyHostToDevice, b->stream));
    }
    sync_tables(b);
    std::vector<int> first(nc);
    for (int i = 0; i < nc; i++) first[i] = b->cases[i].red_first ? 1 : 0;
    HIPX(hipMemcpyAsync(b->dfirst, first.data(), sizeof(int) * nc, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(batch_residual_partials_kernel<kResThreads>, dim3(kResBlocksPerCase * nc),
                       dim3(kResThreads), 0, b->stream, b->dtab, b->duprev, b->dfirst, b->dpart);
    HIPX(hipGetLastError());
    hipLaunchKernelGGL(residual_final_kernel<kResThreads>, dim3(nc), dim3(kResThreads), 0, b->stream, b->dpart,
                       kResBlocksPerCase, b->dres);
    HIPX(hipGetLastError());
    std::vector<double> r(3 * (size_t)nc);
    HIPX(hipMemcpyAsync(r.data(), b->dres, sizeof(double) * r.size(), hipMemcpyDeviceToHost, b->stream));
    HIPX(hipStreamSynchronize(b->stream));
    ke.resize(nc);
    res.resize(nc);
    for (int i = 0; i < nc; i++) {
        ke[i] = r[3 * i];
        const double du = r[3 * i + 1], uu = r[3 * i + 2];
        res[i] = uu > 0 ? std::sqrt(du / uu) : (du > 0 ? INFINITY : 0.0);
        b->cases[i].red_first = false;
    }
}

// Builds a batch with fixed tiling (no tuning).  Throws on failure.
void build(lbm_batch *b, const lbm_batch_case *cs, int ncases) {
    const lbm_batch_config &cfg = b->cfg;
    b->Q = cfg.velocity_set;
    b->TX = cfg.tile_x;
    b->TY = cfg.tile_y;
    b->LZT = cfg.tile_z;
    b->persist = cfg.persistent == 1;
    if (!batch_supported(b->Q, b->TX, b->TY, b->persist)) throw ArgError{LBM_ERR_ARG, "unsupported batch tiling"};
    b->cases.resize(ncases);
    int64_t total_items = 0;
    for (int i = 0; i < ncases; i++) {
        Case &c = b->cases[i];
        c.c = cs[i];
        const auto &k = c.c;
        const std::string tag = "case " + std::to_string(i) + ": ";
        if (k.nx < 2 || k.ny < 2 || k.nz < 2) throw ArgError{LBM_ERR_SIZE, tag + "mesh dimensions must be >= 2"};
        if (!(k.Re > 0) || !std::isfinite(k.u_inf)) throw ArgError{LBM_ERR_ARG, tag + "Re must be > 0, u_inf finite"};
        if (k.nx % b->TX || k.ny % b->TY)
            throw ArgError{LBM_ERR_SIZE, tag + "nx and ny must be multiples of the tile footprint " +
                                             std::to_string(b->TX) + "x" + std::to_string(b->TY)};
        c.n = (int64_t)k.nx * k.ny * k.nz;
        if (c.n > ((int64_t)1 << 30))
            throw ArgError{LBM_ERR_SIZE, tag + "mesh too large: at most 2^30 nodes (4 GiB per moment array)"};
        c.lz = lz_for(k.nz, b->LZT);
        c.nbx = k.nx / b->TX;
        c.nby = k.ny / b->TY;
        c.nc = k.nz / c.lz;
        relaxation(k.Re, k.u_inf, k.nx, &c.omega, &c.tau);
        c.fwords = face_words(b->Q, k.nx, k.ny, k.nz, b->TX, b->TY, c.lz, &c.yoff, &c.zoff);
        if (c.fwords >= (int64_t)INT32_MAX) throw ArgError{LBM_ERR_SIZE, tag + "face buffers exceed 2^31 words"};
        total_items += c.grid();
    }
    if (total_items >= (int64_t)INT32_MAX) throw ArgError{LBM_ERR_SIZE, "too many work items"};
    if (cfg.stream)
        b->stream = (hipStream_t)cfg.stream;
    else {
        HIPX(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->own_stream = true;
    }
    for (auto &c : b->cases) {
        HIPX(hipMalloc(&c.mom, sizeof(float) * 10 * c.n));
        HIPX(hipMemsetAsync(c.mom, 0, sizeof(float) * 10 * c.n, b->stream));
        HIPX(hipMalloc(&c.faces, sizeof(float) * 2 * c.fwords));
        HIPX(hipMemsetAsync(c.faces, 0, sizeof(float) * 2 * c.fwords, b->stream));
    }
    HIPX(hipMalloc(&b->dtab, sizeof(StepParams) * 2 * ncases));
    HIPX(hipMalloc(&b->ditems, sizeof(BatchItem) * std::max<int64_t>(total_items, 1)));
    if (b->persist) {
        int dev = 0, cus = 0, nb = 0;
        HIPX(hipGetDevice(&dev));
        HIPX(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        (void)dispatch<OccF>(b->Q, b->TX, b->TY, true, &nb);
        b->grid_persist = std::max(1, nb) * cus;
    }
    for (int i = 0; i < ncases; i++) fill_faces(b, i);
    sync_tables(b);
    HIPX(hipStreamSynchronize(b->stream));
}

lbm_batch *create_fixed(const lbm_batch_config &cfg, const lbm_batch_case *cs, int ncases, int *code,
                        std::string *msg) {
    lbm_batch *b = new (std::nothrow) lbm_batch();
    if (!b) {
        *code = LBM_ERR_NOMEM;
        *msg = "host allocation failed";
        return nullptr;
    }
    b->cfg = cfg;
    try {
        HIPX(hipSetDevice(cfg.device));
        build(b, cs, ncases);
        *code = LBM_OK;
        return b;
    } catch (const HipError &e) {
        *code = LBM_ERR_HIP;
        *msg = std::string("HIP error ") + hipGetErrorString(e.e) + " in " + e.what;
    } catch (const ArgError &e) {
        *code = e.code;
        *msg = e.msg;
    } catch (const std::bad_alloc &) {
        *code = LBM_ERR_NOMEM;
        *msg = "host allocation failed";
    }
    b->release();
    delete b;
    return nullptr;
}

// ms per step of a built batch: 2 untimed steps, then `steps` timed ones.
double time_batch(lbm_batch *b, int steps) {
    do_steps(b, 2);
    prepare_graphs(b, steps);
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, b->stream);
    if (e == hipSuccess) {
        try {
            do_steps(b, steps);
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        e = hipEventRecord(e1, b->stream);
    }
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPX(e);
    return ms / steps;
}

const int kTiles[][2] = {{32, 16}, {16, 16}, {32, 8}, {8, 8}};
// single-case footprints: also the 1024-thread ones of the single-case kernel
// (lbm_create rejects the combinations a velocity set does not build)
const int kCaseTiles[][2] = {{32, 32}, {64, 16}, {32, 16}, {16, 16}, {32, 8}, {8, 8}};
const int kLzTargets[] = {8, 16, 32, 64};

}  // namespace

extern "C" {

int lbm_batch_create(const lbm_batch_config *cfg, const lbm_batch_case *cases, int ncases, lbm_batch **out) {
    g_batch_create_error.clear();
    if (!cfg || !cases || !out || ncases <= 0) {
        g_batch_create_error = "null argument or no cases";
        return LBM_ERR_ARG;
    }
    *out = nullptr;
    if (cfg->velocity_set != 19 && cfg->velocity_set != 27) {
        g_batch_create_error = "velocity_set must be 19 or 27";
        return LBM_ERR_ARG;
    }
    if (cfg->persistent < -1 || cfg->persistent > 1 || cfg->tile_z < 0) {
        g_batch_create_error = "bad persistent or tile_z";
        return LBM_ERR_ARG;
    }
    const bool tune = cfg->tile_x == 0 || cfg->tile_y == 0 || cfg->tile_z == 0 || cfg->persistent < 0;
    lbm_batch_config best = *cfg;
    std::string report;
    if (tune) {
        double best_ms = INFINITY;
        char line[160];
        for (const auto &t : kTiles) {
            if ((cfg->tile_x && cfg->tile_x != t[0]) || (cfg->tile_y && cfg->tile_y != t[1])) continue;
            bool fits = true;
            for (int i = 0; i < ncases; i++) fits = fits && cases[i].nx % t[0] == 0 && cases[i].ny % t[1] == 0;
            if (!fits) continue;
            for (int lzt : kLzTargets) {
                if (cfg->tile_z && cfg->tile_z != lzt) continue;
                for (int pers = 0; pers < 2; pers++) {
                    if (cfg->persistent >= 0 && cfg->persistent != pers) continue;
                    if (!batch_supported(cfg->velocity_set, t[0], t[1], pers == 1)) continue;
                    lbm_batch_config c = *cfg;
                    c.tile_x = t[0];
                    c.tile_y = t[1];
                    c.tile_z = lzt;
                    c.persistent = pers;
                    int code = 0;
                    std::string msg;
                    lbm_batch *b = create_fixed(c, cases, ncases, &code, &msg);
                    if (!b) {
                        std::snprintf(line, sizeof line, "%2dx%-2d lz<=%-2d %-10s  failed: %s\n", t[0], t[1], lzt,
                                      pers ? "persistent" : "per-item", msg.c_str());
                        report += line;
                        if (code == LBM_ERR_HIP) {  // the device is in trouble: stop here
                            g_batch_create_error = msg;
                            return code;
                        }
                        continue;
                    }
                    double ms = INFINITY;
                    const int rc = guarded(b, [&] { ms = time_batch(b, 8); });
                    const std::string berr = b->err;
                    lbm_batch_destroy(b);
                    if (rc != LBM_OK) {
                        g_batch_create_error = "autotune: " + berr;
                        return rc;
                    }
                    std::snprintf(line, sizeof line, "%2dx%-2d lz<=%-2d %-10s  %.4f ms/step\n", t[0], t[1], lzt,
                                  pers ? "persistent" : "per-item", ms);
                    report += line;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = c;
                    }
                }
            }
        }
        if (!(best_ms < INFINITY)) {
            g_batch_create_error = "autotune: no tiling fits every case\n" + report;
            return LBM_ERR_SIZE;
        }
    }
    int code = 0;
    std::string msg;
    lbm_batch *b = create_fixed(best, cases, ncases, &code, &msg);
    if (!b) {
        g_batch_create_error = msg;
        return code;
    }
    b->report = report;
    *out = b;
    return LBM_OK;
}

const char *lbm_batch_create_error(void) { return g_batch_create_error.c_str(); }

int lbm_batch_destroy(lbm_batch *b) {
    if (!b) return LBM_ERR_ARG;
    (void)hipSetDevice(b->cfg.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->release();
    delete b;
    return LBM_OK;
}

const char *lbm_batch_last_error(const lbm_batch *b) { return b ? b->err.c_str() : "null batch"; }

int lbm_batch_get_tiling(const lbm_batch *b, int *tx, int *ty, int *tz, int *persistent) {
    if (!b) return LBM_ERR_ARG;
    if (tx) *tx = b->TX;
    if (ty) *ty = b->TY;
    if (tz) *tz = b->LZT;
    if (persistent) *persistent = b->persist ? 1 : 0;
    return LBM_OK;
}

int lbm_batch_tuning_report(const lbm_batch *b, char *buf, int cap) {
    if (!b || !buf || cap <= 0) return LBM_ERR_ARG;
    const size_t k = std::min(b->report.size(), (size_t)cap - 1);
    std::memcpy(buf, b->report.data(), k);
    buf[k] = 0;
    return LBM_OK;
}

int lbm_batch_work_items(const lbm_batch *b, int *items) {
    if (!b || !items) return LBM_ERR_ARG;
    int n = 0;
    for (const auto &c : b->cases)
        if (c.active) n += c.grid();
    *items = n;
    return LBM_OK;
}

int lbm_batch_set_active(lbm_batch *b, const uint8_t *active) {
    return guarded(b, [&] {
        if (!active) throw ArgError{LBM_ERR_ARG, "null active array"};
        HIPX(hipStreamSynchronize(b->stream));
        bool changed = false;
        for (int i = 0; i < b->ncases(); i++) {
            const bool a = active[i] != 0;
            changed = changed || a != b->cases[i].active;
            b->cases[i].active = a;
        }
        if (changed) {
            b->tables_dirty = true;
            b->drop_graphs();
        }
    });
}

int lbm_batch_step(lbm_batch *b, int n) {
    return guarded(b, [&] {
        if (n < 0) throw ArgError{LBM_ERR_ARG, "negative step count"};
        do_steps(b, n);
    });
}

int lbm_batch_synchronize(lbm_batch *b) {
    return guarded(b, [&] { HIPX(hipStreamSynchronize(b->stream)); });
}

int lbm_batch_step_timed(lbm_batch *b, int n, double *ms) {
    return guarded(b, [&] {
        if (n < 0 || !ms) throw ArgError{LBM_ERR_ARG, "bad argument"};
        prepare_graphs(b, n);  // tables and graphs outside the timed region
        hipEvent_t e0, e1;
        HIPX(hipEventCreate(&e0));
        HIPX(hipEventCreate(&e1));
        hipError_t e = hipEventRecord(e0, b->stream);
        if (e == hipSuccess) {
            try {
                do_steps(b, n);
            } catch (...) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                throw;
            }
            e = hipEventRecord(e1, b->stream);
        }
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPX(e);
        *ms = f;
    });
}

int lbm_batch_get_steps(const lbm_batch *b, int64_t *steps) {
    if (!b || !steps) return LBM_ERR_ARG;
    for (int i = 0; i < b->ncases(); i++) steps[i] = b->cases[i].t;
    return LBM_OK;
}

int lbm_batch_set_moments(lbm_batch *b, int cs, const float *const m[10], int layout) {
    return guarded(b, [&] {
        check_case(b, cs);
        if (!m) throw ArgError{LBM_ERR_ARG, "null moment array"};
        Case &c = b->cases[cs];
        check_layout(c, layout);
        HIPX(hipStreamSynchronize(b->stream));
        const auto perm = make_perm(layout, c.c.nx, c.c.ny, c.c.nz, b->TX, b->TY);
        std::vector<float> tmp((size_t)c.n);
        for (int k = 0; k < 10; k++) {
            if (!m[k]) throw ArgError{LBM_ERR_ARG, "null moment array"};
            for (int64_t u = 0; u < c.n; u++) tmp[(size_t)perm[(size_t)u]] = m[k][u];
            c.h2d(k, tmp.data(), b->stream);
            HIPX(hipStreamSynchronize(b->stream));
        }
        fill_faces(b, cs);
        HIPX(hipStreamSynchronize(b->stream));
    });
}

int lbm_batch_get_moments(lbm_batch *b, int cs, float *m[10], int layout) {
    return guarded(b, [&] {
        check_case(b, cs);
        if (!m) throw ArgError{LBM_ERR_ARG, "null moment array"};
        Case &c = b->cases[cs];
        check_layout(c, layout);
        HIPX(hipStreamSynchronize(b->stream));
        const auto perm = make_perm(layout, c.c.nx, c.c.ny, c.c.nz, b->TX, b->TY);
        std::vector<float> tmp((size_t)c.n);
        for (int k = 0; k < 10; k++) {
            if (!m[k]) throw ArgError{LBM_ERR_ARG, "null moment array"};
            c.d2h(k, tmp.data(), b->stream);
            HIPX(hipStreamSynchronize(b->stream));
            for (int64_t u = 0; u < c.n; u++) m[k][u] = tmp[(size_t)perm[(size_t)u]];
        }
    });
}

int lbm_batch_reduce(lbm_batch *b, double *ke, double *res) {
    return guarded(b, [&] {
        std::vector<double> k, r;
        reduce_all(b, k, r);
        for (int i = 0; i < b->ncases(); i++) {
            if (ke) ke[i] = k[i];
            if (res) res[i] = r[i];
        }
    });
}

int lbm_batch_run(lbm_batch *b, double tol, int check, int64_t max_steps, int64_t *steps, double *residual,
                  double *ms) {
    return guarded(b, [&] {
        if (check <= 0 || max_steps < 0 || !(tol >= 0)) throw ArgError{LBM_ERR_ARG, "bad run parameters"};
        const auto t0 = std::chrono::steady_clock::now();
        const int nc = b->ncases();
        std::vector<double> ke, res(nc, INFINITY);
        std::vector<uint8_t> act(nc);
        for (int i = 0; i < nc; i++) {
            Case &c = b->cases[i];
            act[i] = c.active && c.t < max_steps;
        }
        auto apply = [&] {
            bool changed = false;
            for (int i = 0; i < nc; i++) {
                changed = changed || (act[i] != 0) != b->cases[i].active;
                b->cases[i].active = act[i] != 0;
            }
            if (changed) {
                HIPX(hipStreamSynchronize(b->stream));
                b->tables_dirty = true;
                b->drop_graphs();
            }
        };
        for (;;) {
            apply();
            bool any = false;
            int64_t n = check;
            for (int i = 0; i < nc; i++)
                if (act[i]) {
                    any = true;
                    n = std::min<int64_t>(n, max_steps - b->cases[i].t);
                }
            if (!any) break;
            do_steps(b, (int)n);
            std::vector<double> r;
            reduce_all(b, ke, r);
            for (int i = 0; i < nc; i++)
                if (act[i]) {
                    res[i] = r[i];
                    if (r[i] < tol || b->cases[i].t >= max_steps) act[i] = 0;
                }
        }
        HIPX(hipStreamSynchronize(b->stream));
        const auto t1 = std::chrono::steady_clock::now();
        for (int i = 0; i < nc; i++) {
            if (steps) steps[i] = b->cases[i].t;
            if (residual) residual[i] = res[i];
        }
        if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    });
}

int lbm_autotune(const lbm_config *cfg, int steps, int *tile_x, int *tile_y, int *tile_z, double *ms_per_step) {
    g_batch_create_error.clear();
    if (!cfg || steps <= 0) return LBM_ERR_ARG;
    double best = INFINITY;
    int bt[3] = {0, 0, 0};
    for (const auto &t : kCaseTiles) {
        if (cfg->nx % t[0] || cfg->ny % t[1]) continue;
        std::vector<int> lzs;
        for (int lzt : kLzTargets) {
            const int lz = lz_for(cfg->nz, lzt);
            if (std::find(lzs.begin(), lzs.end(), lz) == lzs.end()) lzs.push_back(lz);
        }
        for (int lz : lzs) {
            lbm_config c = *cfg;
            c.tile_x = t[0];
            c.tile_y = t[1];
            c.tile_z = lz;
            lbm_handle *h = nullptr;
            const int rc = lbm_create(&c, &h);
            if (rc == LBM_ERR_HIP) {
                g_batch_create_error = lbm_create_error();
                return rc;
            }
            if (rc != LBM_OK) continue;
            double ms = 0;
            int r = lbm_step(h, 2);
            if (r == LBM_OK) r = lbm_step_timed(h, steps, &ms);
            if (r != LBM_OK) {
                g_batch_create_error = lbm_last_error(h);
                lbm_destroy(h);
                return r;
            }
            lbm_destroy(h);
            if (ms / steps < best) {
                best = ms / steps;
                bt[0] = t[0];
                bt[1] = t[1];
                bt[2] = lz;
            }
        }
    }
    if (!(best < INFINITY)) {
        g_batch_create_error = "no supported tiling fits the mesh";
        return LBM_ERR_SIZE;
    }
    if (tile_x) *tile_x = bt[0];
    if (tile_y) *tile_y = bt[1];
    if (tile_z) *tile_z = bt[2];
    if (ms_per_step) *ms_per_step = best;
    return LBM_OK;
}

}  // extern "C"
