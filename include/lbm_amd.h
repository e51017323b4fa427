/*
 * lbm_amd.h -- C ABI of the MI355X moment-representation LBM solver.
 *
 * The reference (geoenergiaUDESC/cudaLBM) exposes this path through a CUDA
 * kernel signature, a CLI and its output files, not through an FFI:
 *   - kernel   applications/momentBasedD3Q19/momentBasedD3Q19.cuh:88-91
 *              applications/momentBasedD3Q27/momentBasedD3Q27.cuh:88-91
 *   - driver   applications/momentBasedD3Q19/momentBasedD3Q19.cu:56-140
 *   - constants uploaded by host::latticeMesh (src/latticeMesh/latticeMesh.cuh:178-209)
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative lbm_status on error;
 *     lbm_last_error() returns the message of the last failure on a handle
 *     (the reference exits the process instead: src/errorHandler.cuh:88-168).
 *   - Moments are the reference's stored values, in this order:
 *     rho-1, 3u, 3v, 3w, 4.5(Pxx/rho-cs2), 9Pxy/rho, 9Pxz/rho,
 *     4.5(Pyy/rho-cs2), 9Pyz/rho, 4.5(Pzz/rho-cs2)
 *     (src/globalFunctions.cuh:482-503, src/velocitySet/velocitySet.cuh:130-142).
 *   - Host buffers are owned by the caller; all device memory is owned by the
 *     handle.  One handle = one HIP stream; calls on a handle must not overlap.
 *   - No C++ exception crosses this boundary.
 */
#ifndef LBM_AMD_H
#define LBM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_handle lbm_handle;

typedef enum {
    LBM_OK = 0,
    LBM_ERR_ARG = -1,      /* bad argument or configuration */
    LBM_ERR_SIZE = -49,    /* mesh not a multiple of the tile (reference ERR_SIZE) */
    LBM_ERR_HIP = -2,      /* HIP runtime failure */
    LBM_ERR_NOMEM = -3,    /* device or host allocation failed */
    LBM_ERR_STATE = -4     /* call not valid in the handle's current state */
} lbm_status;

typedef enum { LBM_STORE_FP32 = 0, LBM_STORE_FP16 = 1 } lbm_storage;

/* Host array layouts accepted by lbm_set_moments / lbm_get_moments. */
typedef enum {
    LBM_LAYOUT_GLOBAL = 0, /* idx = x + nx*(y + ny*z) */
    LBM_LAYOUT_TILED8 = 1  /* reference 8^3 block order (src/globalFunctions.cuh:347-370),
                              the order .LBMBin files carry */
} lbm_layout;

/* Function-object selection bits (src/functionObjects/functionObjects.cuh:65-70). */
enum {
    LBM_FO_K = 1u << 0,     /* k, instantaneous (FP32) */
    LBM_FO_KMEAN = 1u << 1, /* kMean, running mean (FP64 accumulator) */
    LBM_FO_S = 1u << 2,     /* S_xx S_xy S_xz S_yy S_yz S_zz, instantaneous (FP32) */
    LBM_FO_SMEAN = 1u << 3, /* SMean, running means (FP64 accumulators) */
    LBM_FO_MOMENTSMEAN = 1u << 4 /* momentsMean: running means of the 10 stored moments
                                    (FP64 accumulators; src/functionObjects/moments.cuh:70-101) */
};

/* Which FO field to read back with lbm_fo_get. */
typedef enum {
    LBM_FIELD_K = 0,
    LBM_FIELD_KMEAN = 1,
    LBM_FIELD_S = 2,
    LBM_FIELD_SMEAN = 3,
    LBM_FIELD_MOMENTSMEAN = 4
} lbm_fo_field;

typedef struct lbm_config {
    int velocity_set; /* 19 or 27 */
    int storage;      /* lbm_storage */
    int nx, ny, nz;
    double Re;        /* omega = 1/(0.5 + 3 u_inf (nx-1)/Re), src/latticeMesh/latticeMesh.cuh:178-182 */
    double u_inf;     /* lid speed: U_North = {u_inf,0,0}, src/latticeMesh/latticeMesh.cuh:183 */
    int device;       /* HIP device ordinal (reference: -GPU <id>) */
    void *stream;     /* hipStream_t to run on, or NULL for a handle-owned stream */
    int tile_x, tile_y; /* workgroup footprint; 0 = choose automatically */
    int tile_z;         /* planes marched per workgroup; 0 = choose automatically */
} lbm_config;

/* Replaces host::latticeMesh + device::array + device::halo construction
 * (momentBasedD3Q19.cu:66-98).  Allocates the moments (initialised to rest)
 * and both face-buffer sets.  At most 2^30 nodes (the kernels address each
 * moment array with 32-bit byte offsets); LBM_ERR_SIZE otherwise. */
int lbm_create(const lbm_config *cfg, lbm_handle **out);
int lbm_destroy(lbm_handle *h);
const char *lbm_last_error(const lbm_handle *h);
/* Message of the last failure of lbm_create (no handle exists then). */
const char *lbm_create_error(void);

/* Geometry in use (chosen automatically unless set in lbm_config; with an
 * automatic tiling lbm_fo_enable may change it) and derived constants. */
int lbm_get_tiling(const lbm_handle *h, int *tile_x, int *tile_y, int *tile_z);
int lbm_get_omega(const lbm_handle *h, double *omega, double *tau);

/* Replaces the host::array initial-condition/restart upload (hostArray.cuh:264-296,
 * memory.cuh:274-282) and haloFace::initialise_pop (haloFace.cuh:244-449): uploads
 * the 10 moment arrays and rebuilds the face buffers from them on the device. */
int lbm_set_moments(lbm_handle *h, const float *const m[10], int layout);
/* Replaces host::toHost(devPtrs) (memory.cuh:157-195). */
int lbm_get_moments(lbm_handle *h, float *m[10], int layout);
/* FP64 upload/download variants (values are rounded to the storage type). */
int lbm_set_moments_f64(lbm_handle *h, const double *const m[10], int layout);
int lbm_get_moments_f64(lbm_handle *h, double *m[10], int layout);

/* Replaces the per-step kernel launch + blockHalo.swap() + runTimeObjects.calculate()
 * of the driver loop (momentBasedD3Q19.cu:126-136).  Advances n steps; launches are
 * recorded once into a hipGraph per (n, face parity) and replayed. Asynchronous. */
int lbm_step(lbm_handle *h, int n);
int lbm_synchronize(lbm_handle *h);
/* Timed variant: hipEvents on the handle's stream around n steps; returns the
 * elapsed milliseconds (includes only device time of the n steps). */
int lbm_step_timed(lbm_handle *h, int n, double *ms);
/* Number of steps taken since creation / the last lbm_set_time. */
int lbm_get_time(const lbm_handle *h, int64_t *t);
int lbm_set_time(lbm_handle *h, int64_t t);

/* Replaces objectRegistry (objectRegistry.cuh:68-117), whose calculate() runs
 * after every step (momentBasedD3Q19.cu:133).  The time means (kMean, SMean,
 * momentsMean) are fused into the step: FP64 accumulators updated from the stored
 * post-collision moments with invNewCount = 1/(t+1) of the absolute step t
 * (kineticEnergy.cuh:249).  The instantaneous fields (k, S) are functions of
 * the current moments alone, so they are formed from them when read, with the
 * same formulas (kineticEnergy.cuh:71-85, strainRateTensor.cuh:71-84).
 * Enabling resets the means to zero.  Synchronizes the handle's stream.  Turning
 * the means on or off may move the moment arrays inside device memory, keeping
 * their values: with an automatic tiling (tile fields 0 at lbm_create) the case
 * is re-tiled for the step with or without the means when its fastest
 * footprint differs (D3Q19 FP32: 32x16 without, 32x32 with the means; D3Q27:
 * 32x16 or 32x32 with one up-going LDS bank without, 16x16 or 32x8 with two
 * banks with the means), otherwise only the arrays' spacing changes.  The
 * move briefly holds the moments (and, re-tiling, the face buffers) twice; when
 * that memory is not available the case stays as it is.  Results do not depend
 * on the tiling (bit for bit).  LBM_ERR_NOMEM when an FO array cannot be
 * allocated. */
int lbm_fo_enable(lbm_handle *h, unsigned mask);
/* Reads an FO field.  LBM_FIELD_K/KMEAN: n values; S/SMEAN: 6*n values in the
 * order xx,xy,xz,yy,yz,zz, component-major; MOMENTSMEAN: 10*n values, the means
 * of the stored moments in the order of lbm_set_moments (rhoMean is the mean of
 * rho-1, as the reference averages the stored arrays), component-major. */
int lbm_fo_get(lbm_handle *h, int field, double *out, int layout);

/* Global FP64 reductions (no reference counterpart): kinetic energy
 * E = sum_nodes 0.5*rho*|u|^2 (physical u) and the velocity residual
 * r = ||u_t - u_prev||_2 / ||u_t||_2 against the velocity field stored by the
 * previous call (r = 1 on the first call).  Deterministic (fixed-order). */
int lbm_reduce(lbm_handle *h, double *kinetic_energy, double *residual);

/* No counterpart in the reference.  lbm_reduce for callers that cannot
 * allocate device memory or synchronize: code under HIP graph capture and
 * tight solve loops.  It allocates nothing (no hipMalloc/hipFree, no
 * allocating library call) and only enqueues kernels on the handle's stream.
 *   uprev        device buffer of uprev_count floats, owned by the caller,
 *                at least 3 * nodes: the previous velocity field in the
 *                library's internal order; read (unless `first`) and then
 *                overwritten with the current field.  lbm_fo_enable may change
 *                that order: pass first = 1 on the next call after it.
 *   first        nonzero: no previous field yet; the residual is reported as 1
 *                (as on lbm_reduce's first call).
 *   out          device memory for 2 doubles: kinetic energy, residual.
 * The per-block partial sums live in a fixed static device buffer of the
 * library (1024 blocks x 3 doubles + 3, 24 KiB per GPU, kept until the
 * process ends), shared by every handle on that GPU: calls that overlap on
 * one GPU (two streams, or two threads) race and must be serialized by the
 * caller.  LBM_ERR_SIZE if uprev_count < 3 * nodes.  Results equal
 * lbm_reduce's bit for bit (same blocks, same summation order). */
int lbm_reduce_noalloc(lbm_handle *h, float *uprev, int64_t uprev_count, int first, double *out);

/* Integer probes of the streaming/halo machinery (test hooks; no reference
 * counterpart).  src_out[d*Q+q] receives 1 + 32*s + q where s is the global
 * x-fastest index of the node whose population q arrives at node d, or 0 when
 * that population is unknown (source outside the domain).  type_out[d] is the
 * node-type bitmask of src/boundaryConditions/normalVector.cuh:321-334. */
int lbm_probe_streaming(lbm_handle *h, uint32_t *src_out, uint8_t *type_out);

/* Tile autotuner (no reference counterpart: the reference launches fixed 8x8x8
 * blocks, momentBasedD3Q19.cu:129 with src/latticeMesh/latticeMesh.cuh:295).
 * Creates a solver for cfg's case with every supported footprint and plane
 * count in turn, times `steps` steps of each and returns the fastest.  cfg's
 * tile fields are ignored. */
int lbm_autotune(const lbm_config *cfg, int steps, int *tile_x, int *tile_y, int *tile_z, double *ms_per_step);

/* ------------------------------------------------------------------ batches
 * Many independent cavities of uneven size, Re and lid speed advancing
 * together: one kernel launch per step covers a tile-to-case table of
 * (case, column, chunk) work items, and steps are replayed from hipGraphs.
 * No reference counterpart: the reference runs one case per process
 * (momentBasedD3Q19.cu:54-137); each case follows the same physics as
 * lbm_create/lbm_step.  Storage is FP32. */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_case {
    int nx, ny, nz;
    double Re, u_inf; /* as in lbm_config */
} lbm_batch_case;

typedef struct lbm_batch_config {
    int velocity_set;   /* 19 or 27 */
    int device;         /* HIP device ordinal */
    void *stream;       /* hipStream_t, or NULL for a batch-owned stream */
    int tile_x, tile_y; /* footprint shared by all cases; 0 = autotune */
    int tile_z;         /* planes per work item: the largest divisor of each case's
                           nz not above tile_z; 0 = autotune */
    int persistent;     /* 0: one workgroup per work item; 1: a persistent grid
                           striding over the items; -1 = autotune */
} lbm_batch_config;

/* Allocates every case at rest.  With any autotune field set, each candidate
 * (footprint x plane count x grid kind) is built and timed on the real batch
 * first; lbm_batch_tuning_report lists them. */
int lbm_batch_create(const lbm_batch_config *cfg, const lbm_batch_case *cases, int ncases, lbm_batch **out);
/* Message of the last failure of lbm_batch_create or lbm_autotune (thread-local). */
const char *lbm_batch_create_error(void);
int lbm_batch_destroy(lbm_batch *b);
const char *lbm_batch_last_error(const lbm_batch *b);
int lbm_batch_get_tiling(const lbm_batch *b, int *tile_x, int *tile_y, int *tile_z, int *persistent);
/* Text table of the autotuner's candidates and their ms per step (empty when
 * nothing was tuned).  Writes at most cap bytes including the NUL. */
int lbm_batch_tuning_report(const lbm_batch *b, char *buf, int cap);
/* Work items of the active cases (the grid of one step when not persistent). */
int lbm_batch_work_items(const lbm_batch *b, int *items);
/* active[i] != 0 keeps case i stepping; an inactive case keeps its fields and
 * step count.  All cases start active. */
int lbm_batch_set_active(lbm_batch *b, const uint8_t *active);
/* Advances every active case n steps (asynchronous). */
int lbm_batch_step(lbm_batch *b, int n);
int lbm_batch_step_timed(lbm_batch *b, int n, double *ms);
int lbm_batch_synchronize(lbm_batch *b);
/* Steps taken by each case (ncases values). */
int lbm_batch_get_steps(const lbm_batch *b, int64_t *steps);
/* Moments of case cs, as lbm_set_moments / lbm_get_moments. */
int lbm_batch_set_moments(lbm_batch *b, int cs, const float *const m[10], int layout);
int lbm_batch_get_moments(lbm_batch *b, int cs, float *m[10], int layout);
/* lbm_reduce for every case at once (ncases values each; either may be NULL). */
int lbm_batch_reduce(lbm_batch *b, double *kinetic_energy, double *residual);
/* Runs the sweep: repeatedly advances the active cases `check` steps (fewer
 * when a case would pass max_steps), reduces, and deactivates each case whose
 * residual is below tol or that reached max_steps, until none is active.
 * steps/residual (ncases each, may be NULL) receive each case's final values;
 * ms the wall time of the whole run. */
int lbm_batch_run(lbm_batch *b, double tol, int check, int64_t max_steps, int64_t *steps, double *residual,
                  double *ms);

#ifdef __cplusplus
}
#endif

#endif
