/*
 * lbm_surrogate.h -- C ABI of the learned warm start: a 4-level 3D
 * convolutional U-Net (channels 16-32-64-128, 3x3x3 kernels) that maps a
 * cubic grid of inputs (lid mask, Re and N as constant channels) to the
 * normalised steady fields u/U, v/U, w/U, p/U^2 of a lid-driven cavity.
 *
 * No reference counterpart: the reference always starts from the
 * initialConditions file (hostArray.cuh:304-359).  The network's forward and
 * backward passes run on hand-written BF16 MFMA kernels with FP32
 * accumulation (FP32 master weights, Adam); everything stays on the device.
 *
 * Layouts: grids are S^3 with S a multiple of 8 (S = 32 for the trained
 * network), x fastest then y then z, channels last:
 *   inputs  [n][S^3][3]  (lid mask, Re code, N code; see lbm_sur_encode)
 *   fields  [n][S^3][4]  (u/U, v/U, w/U, p/U^2 with p = (rho-1)/3 minus its
 *                         domain mean: the cavity's mean density drifts with
 *                         the elapsed steps and is not predicted)
 * Every function returns 0 or a negative lbm_status (include/lbm_amd.h).
 */
#ifndef LBM_SURROGATE_H
#define LBM_SURROGATE_H

#include <stdint.h>

#include "lbm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_sur lbm_sur;

/* Network for S^3 grids, weights drawn from `seed` (He-normal, zero bias). */
int lbm_sur_create(int device, int S, uint64_t seed, lbm_sur **out);
int lbm_sur_destroy(lbm_sur *s);
const char *lbm_sur_last_error(const lbm_sur *s);
const char *lbm_sur_create_error(void);

/* Number of FP32 master parameters (all weights then all biases). */
int64_t lbm_sur_num_params(const lbm_sur *s);
int lbm_sur_get_params(lbm_sur *s, float *out);
int lbm_sur_set_params(lbm_sur *s, const float *in);

/* The three input channels of one case: lid mask (1 on the y = S-1 plane),
 * c_Re = ln(Re/800)/ln(4), c_N = log2(N/64). */
int lbm_sur_encode(int S, double Re, int N, float *inputs);

/* Loss and gradient of mean((pred - target)^2) over n samples, without
 * updating (for checks); grad receives lbm_sur_num_params values. */
int lbm_sur_gradient(lbm_sur *s, int n, const float *inputs, const float *targets, double *loss, float *grad);

/* Keeps n training samples resident on the device. */
int lbm_sur_load_data(lbm_sur *s, int n, const float *inputs, const float *targets);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over the resident samples in
 * shuffled mini-batches of `batch`, `epochs` times; loss_per_epoch (may be
 * NULL) receives each epoch's mean loss. */
int lbm_sur_train(lbm_sur *s, int epochs, int batch, float lr, uint64_t seed, double *loss_per_epoch);

/* Forward pass on n samples. */
int lbm_sur_predict(lbm_sur *s, int n, const float *inputs, float *fields);

/* Warm start of a solver: predicts the normalised fields for (Re, N = h's nx)
 * on the network's grid, interpolates them trilinearly onto the solver's
 * nodes (network node i sits at i (N-1)/(S-1)), converts to stored moments at
 * equilibrium (rho-1 = 3 p U^2, 3u, 3v, 3w, 4.5 u^2, 9 uv, 9 uw, 4.5 v^2,
 * 9 vw, 4.5 w^2 with u = U u/U) and uploads them (lbm_set_moments).  The
 * solver must be cubic. */
int lbm_sur_warm_start(lbm_sur *s, lbm_handle *h, double Re, double u_inf);

typedef struct lbm_solve_info {
    int64_t steps;        /* steps taken, including any discarded warm-start steps */
    double residual;      /* last residual */
    double ms;            /* wall time: prediction, conversion, upload and stepping */
    double ms_warm;       /* part of ms spent on prediction, conversion and upload */
    double first_residual; /* residual at the first check */
    int warm;             /* 1 if the run kept the warm start */
    int fell_back;        /* 1 if the warm start was rejected (NaN, or first residual >= 1) */
} lbm_solve_info;

/* Runs a solver (at its current state when s is NULL, e.g. rest) until the
 * velocity residual r = ||u_t - u_{t-check}|| / ||u_t|| (FP64, lbm_reduce)
 * drops below tol or max_steps steps were taken.  With s, the solver is first
 * warm-started from the network; the warm start is tested at the first check:
 * a NaN or a residual >= 1 (what a start from rest shows at its first check,
 * since u_0 = 0) rejects it, the moments are reset to rest and the run
 * continues from there (the discarded steps still count in steps and ms). */
int lbm_solve_to_residual(lbm_handle *h, lbm_sur *s, double Re, double u_inf, double tol, int check,
                          int64_t max_steps, lbm_solve_info *info);

#ifdef __cplusplus
}
#endif

#endif
