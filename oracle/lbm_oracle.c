/*
 * lbm_oracle.c -- FP64 scalar CPU restatement of the reference LBM step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP solver and the timed CPU
 * baseline.  The product (cudalbm_amd/, include/lbm_amd.h) never links it.
 *
 * What each part follows (reference file:line, paths relative to the reference
 * root) and what pins it (tests under tests/):
 *   velocity sets ........ D3Q19.cuh:98-288, D3Q27.cuh:98-225
 *                          pinned: weight/isotropy sums (test_oracle_closed_forms)
 *   reconstruct .......... D3Q19.cuh:346-374, D3Q27.cuh:294-331
 *                          pinned: moments(reconstruct(m)) == m round trip
 *   interior moments ..... velocitySet.cuh:398-417 (pinned by the same round trip)
 *   incomingSwitch ....... velocitySet.cuh:204-220
 *                          pinned: known <=> source inside the domain, exhaustive
 *   boundary moments ..... boundaryConditions.cuh:99-554
 *                          pinned: rest fixed point at every node class (exact
 *                          rationals 1 and 125/126); all 26 classes of both
 *                          velocity sets against an independent derivation
 *                          solved by Newton in the tests (known density of the
 *                          post-collision populations, known off-diagonals of
 *                          the pre-collision ones), at several omega and U --
 *                          the lid's NORTH-WEST/EAST edges and corners match it
 *                          with the lid reversed (-U) in the density equation,
 *                          the D3Q27 NORTH-BACK/FRONT edges and lid corners with
 *                          the D3Q19 closure; the reversed-lid equilibrium as a
 *                          closed form; x and z mirror symmetry of the lid
 *                          closures
 *   scale / collide ...... velocitySet.cuh:130-142, secondOrder.cuh:93-105
 *                          pinned: affine closed form + equilibrium fixed point
 *   node type ............ normalVector.cuh:321-334 (pinned: exhaustive counts)
 *   omega ................ latticeMesh.cuh:178-182 (pinned: closed-form values)
 *   initial fields ....... hostArray.cuh:304-359, boundaryValue.cuh:147-254
 *   function objects ..... kineticEnergy.cuh:71-85, strainRateTensor.cuh:71-84,
 *                          functionObjects.cuh:79-83, moments.cuh:70-101
 *                          (pinned: closed forms, running mean of 1..n =
 *                          (n+1)/2; lbmo_run_means against step-by-step
 *                          lbmo_k/lbmo_S/lbmo_time_average and the mean of a
 *                          stationary state)
 *   tile source map ...... streaming.cuh:126-154, halo.cuh:171-293 (load table),
 *                          halo.cuh:813-904 (save slots), globalFunctions.cuh:415-476
 *                          pinned: equals periodic(x - c_q) for every node, and
 *                          x - c_q with no wrap wherever the population is known
 *   whole step ........... momentBasedD3Q19.cuh:117-184: pull streaming f_q(x) =
 *                          f*_q(x - c_q); pinned by mass conservation on closed
 *                          boxes at rest, the rest fixed point, and symmetry of
 *                          the cavity about the z mid-plane.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- tables */

static const int C27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
    {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, 0, 1},   {-1, 0, -1}, {0, 1, 1},
    {0, -1, -1}, {1, -1, 0},  {-1, 1, 0},  {1, 0, -1},  {-1, 0, 1},  {0, 1, -1},
    {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
    {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}};

/* Face slot order, faces x0,x1,y0,y1,z0,z1 (halo.cuh:813-904). */
static const int SLOTS[6][9] = {{2, 8, 10, 14, 16, 20, 22, 24, 25},
                                {1, 7, 9, 13, 15, 19, 21, 23, 26},
                                {4, 8, 12, 13, 18, 20, 22, 23, 26},
                                {3, 7, 11, 14, 17, 19, 21, 24, 25},
                                {6, 10, 12, 15, 17, 20, 21, 24, 26},
                                {5, 9, 11, 16, 18, 19, 22, 23, 25}};

static double weight(int Q, int q) {
    const int s = abs(C27[q][0]) + abs(C27[q][1]) + abs(C27[q][2]);
    if (Q == 19) {
        static const double w19[3] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0};
        return w19[s];
    }
    static const double w27[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};
    return w27[s];
}

int lbmo_velocity_set(int Q, int *cx, int *cy, int *cz, double *w) {
    if (Q != 19 && Q != 27) return -1;
    for (int q = 0; q < Q; q++) {
        if (cx) cx[q] = C27[q][0];
        if (cy) cy[q] = C27[q][1];
        if (cz) cz[q] = C27[q][2];
        if (w) w[q] = weight(Q, q);
    }
    return 0;
}

int lbmo_qf(int Q) { return Q == 19 ? 5 : (Q == 27 ? 9 : -1); }

int lbmo_face_slots(int Q, int *out) {
    const int qf = lbmo_qf(Q);
    if (qf < 0) return -1;
    for (int f = 0; f < 6; f++)
        for (int s = 0; s < qf; s++) out[f * qf + s] = SLOTS[f][s];
    return 0;
}

/* ------------------------------------------------------- pointwise maths */

static const double CS2 = 1.0 / 3.0;

int lbmo_reconstruct(int Q, const double *m, double *f) {
    if (Q != 19 && Q != 27) return -1;
    /* pics2 = 1 - cs2 (mxx + myy + mzz)  (D3Q19.cuh:348) */
    const double pics2 = 1.0 - CS2 * (m[4] + m[7] + m[9]);
    for (int q = 0; q < Q; q++) {
        const double x = C27[q][0], y = C27[q][1], z = C27[q][2];
        const double a = pics2 + x * m[1] + y * m[2] + z * m[3] + x * x * m[4] + y * y * m[7] +
                         z * z * m[9] + x * y * m[5] + x * z * m[6] + y * z * m[8];
        f[q] = m[0] * weight(Q, q) * a;
    }
    return 0;
}

int lbmo_moments_interior(int Q, const double *f, double *m) {
    if (Q != 19 && Q != 27) return -1;
    double s[10] = {0};
    for (int q = 0; q < Q; q++) {
        const double x = C27[q][0], y = C27[q][1], z = C27[q][2];
        s[0] += f[q];
        s[1] += f[q] * x;
        s[2] += f[q] * y;
        s[3] += f[q] * z;
        s[4] += f[q] * x * x;
        s[5] += f[q] * x * y;
        s[6] += f[q] * x * z;
        s[7] += f[q] * y * y;
        s[8] += f[q] * y * z;
        s[9] += f[q] * z * z;
    }
    const double inv = 1.0 / s[0];
    m[0] = s[0];
    m[1] = s[1] * inv;
    m[2] = s[2] * inv;
    m[3] = s[3] * inv;
    m[4] = s[4] * inv - CS2;
    m[5] = s[5] * inv;
    m[6] = s[6] * inv;
    m[7] = s[7] * inv - CS2;
    m[8] = s[8] * inv;
    m[9] = s[9] * inv - CS2;
    return 0;
}

int lbmo_known(int Q, int q, int bits) {
    (void)Q;
    const int x = C27[q][0], y = C27[q][1], z = C27[q][2];
    /* East (bit1): incoming if cx<0; West (bit0): cx>0; North (bit3): cy<0;
     * South (bit2): cy>0; Front (bit5): cz<0; Back (bit4): cz>0. */
    const int cond_x = ((bits & 2) && x < 0) || ((bits & 1) && x > 0);
    const int cond_y = ((bits & 8) && y < 0) || ((bits & 4) && y > 0);
    const int cond_z = ((bits & 32) && z < 0) || ((bits & 16) && z > 0);
    return !(cond_x || cond_y || cond_z);
}

enum {
    W_ = 1, E_ = 2, S_ = 4, N_ = 8, B_ = 16, F_ = 32
};

/* rho for the "static edge" family: sign +1 is the (rho_I - m rho_I + m rho_I w)
 * form, sign -1 the -(-rho_I - m rho_I + m rho_I w) form
 * (boundaryConditions.cuh:181-303). Returns rho and the off-diagonal. */
static void static_edge(int sign, double rho_I, double mI, double w, double *rho, double *mab) {
    if (sign > 0) {
        *rho = 36.0 * (rho_I - mI * rho_I + mI * rho_I * w) / (24.0 + w);
        *mab = (36.0 * mI * rho_I - *rho) / (9.0 * *rho);
    } else {
        *rho = -36.0 * (-rho_I - mI * rho_I + mI * rho_I * w) / (24.0 + w);
        *mab = (36.0 * mI * rho_I + *rho) / (9.0 * *rho);
    }
}

int lbmo_moments_boundary(int Q, const double *f, int bits, double omega, const double *U,
                          double *m) {
    if (Q != 19 && Q != 27) return -1;
    double rho_I = 0, sxy = 0, sxz = 0, syz = 0;
    for (int q = 0; q < Q; q++) {
        if (!lbmo_known(Q, q, bits)) continue;
        const double x = C27[q][0], y = C27[q][1], z = C27[q][2];
        rho_I += f[q];
        sxy += f[q] * x * y;
        sxz += f[q] * x * z;
        syz += f[q] * y * z;
    }
    const double inv_rho_I = 1.0 / rho_I;
    const double mxy_I = sxy * inv_rho_I, mxz_I = sxz * inv_rho_I, myz_I = syz * inv_rho_I;

    /* Face velocities in the reference switch order West, East, North, South,
     * Front, Back (boundaryConditions.cuh:110-124, 543-554). */
    const int sw[6] = {!!(bits & W_), !!(bits & E_), !!(bits & N_),
                       !!(bits & S_), !!(bits & F_), !!(bits & B_)};
    const double nb = sw[0] + sw[1] + sw[2] + sw[3] + sw[4] + sw[5];
    for (int a = 0; a < 3; a++) {
        double s = 0;
        for (int k = 0; k < 6; k++) s += sw[k] * U[3 * k + a];
        m[1 + a] = s / nb;
    }
    m[4] = m[1] * m[1];
    m[7] = 0.0;
    m[9] = 0.0;
    const double w = omega;
    const double u = U[3 * 2 + 0]; /* U_North[0] */
    double rho = 0, mxy = 0, mxz = 0, myz = 0;
    switch (bits & 0x3F) {
    case S_ | W_ | B_:
    case S_ | W_ | F_:
    case S_ | E_ | B_:
    case S_ | E_ | F_:
        rho = (Q == 19) ? 12.0 * rho_I / 7.0 : 216.0 * rho_I / 125.0;
        break;
    case S_ | W_: static_edge(+1, rho_I, mxy_I, w, &rho, &mxy); break;
    case S_ | E_: static_edge(-1, rho_I, mxy_I, w, &rho, &mxy); break;
    case W_ | B_: static_edge(+1, rho_I, mxz_I, w, &rho, &mxz); break;
    case W_ | F_: static_edge(-1, rho_I, mxz_I, w, &rho, &mxz); break;
    case E_ | B_: static_edge(-1, rho_I, mxz_I, w, &rho, &mxz); break;
    case E_ | F_: static_edge(+1, rho_I, mxz_I, w, &rho, &mxz); break;
    case S_ | B_: static_edge(+1, rho_I, myz_I, w, &rho, &myz); break;
    case S_ | F_: static_edge(-1, rho_I, myz_I, w, &rho, &myz); break;
    case W_:
    case E_:
        rho = 6.0 * rho_I / 5.0;
        mxy = 2.0 * mxy_I * rho_I / rho;
        mxz = 2.0 * mxz_I * rho_I / rho;
        break;
    case S_:
        rho = 6.0 * rho_I / 5.0;
        mxy = 2.0 * mxy_I * rho_I / rho;
        myz = 2.0 * myz_I * rho_I / rho;
        break;
    case B_:
    case F_:
        rho = 6.0 * rho_I / 5.0;
        mxz = 2.0 * mxz_I * rho_I / rho;
        myz = 2.0 * myz_I * rho_I / rho;
        break;
    case N_:
        rho = 6.0 * rho_I / 5.0;
        mxy = (6.0 * mxy_I * rho_I - u * rho) / (3.0 * rho);
        myz = 2.0 * myz_I * rho_I / rho;
        break;
    case N_ | W_ | B_:
    case N_ | W_ | F_:
        rho = -24.0 * rho_I / (-14.0 - 8.0 * u + 9.0 * u * u);
        break;
    case N_ | E_ | B_:
    case N_ | E_ | F_:
        rho = -24.0 * rho_I / (-14.0 + 8.0 * u + 9.0 * u * u);
        break;
    case N_ | B_:
        rho = 72.0 * (-rho_I - myz_I * rho_I + myz_I * rho_I * w) /
              (-48.0 - 2.0 * w + 3.0 * u * u * w);
        myz = (72.0 * myz_I * rho_I + 2.0 * rho - 3.0 * u * u * rho) / (18.0 * rho);
        break;
    case N_ | F_:
        rho = -72.0 * (rho_I - myz_I * rho_I + myz_I * rho_I * w) /
              (-48.0 - 2.0 * w + 3.0 * u * u * w);
        myz = (72.0 * myz_I * rho_I - 2.0 * rho + 3.0 * u * u * rho) / (18.0 * rho);
        break;
    case N_ | E_:
        rho = 36.0 * (rho_I - mxy_I * rho_I + mxy_I * rho_I * w) /
              (24.0 - 18.0 * u - 18.0 * u * u + w + 3.0 * u * w + 3.0 * u * u * w);
        mxy = (36.0 * mxy_I * rho_I - rho - 3.0 * u * rho - 3.0 * u * u * rho) / (9.0 * rho);
        break;
    case N_ | W_:
        rho = -36.0 * (-rho_I - mxy_I * rho_I + mxy_I * rho_I * w) /
              (24.0 + 18.0 * u - 18.0 * u * u + w - 3.0 * u * w + 3.0 * u * u * w);
        mxy = (36.0 * mxy_I * rho_I + rho - 3.0 * u * rho + 3.0 * u * u * rho) / (9.0 * rho);
        break;
    default:
        return -2; /* not a boundary node type of a box */
    }
    m[0] = rho;
    m[5] = mxy;
    m[6] = mxz;
    m[8] = myz;
    return 0;
}

void lbmo_scale(double *m) {
    m[1] *= 3.0;
    m[2] *= 3.0;
    m[3] *= 3.0;
    m[4] *= 4.5;
    m[5] *= 9.0;
    m[6] *= 9.0;
    m[7] *= 4.5;
    m[8] *= 9.0;
    m[9] *= 4.5;
}

void lbmo_collide(double *m, double omega) {
    const double t = 1.0 - omega, h = 0.5 * omega;
    m[4] = t * m[4] + h * m[1] * m[1];
    m[7] = t * m[7] + h * m[2] * m[2];
    m[9] = t * m[9] + h * m[3] * m[3];
    m[5] = t * m[5] + omega * m[1] * m[2];
    m[6] = t * m[6] + omega * m[1] * m[3];
    m[8] = t * m[8] + omega * m[2] * m[3];
}

int lbmo_node_type(int x, int y, int z, int nx, int ny, int nz) {
    const int b = (x == 0) | (x == nx - 1) << 1 | (y == 0) << 2 | (y == ny - 1) << 3 |
                  (z == 0) << 4 | (z == nz - 1) << 5;
    return b | ((b != 0) << 6);
}

double lbmo_omega(double u_inf, double Re, int nx) {
    const double nu = u_inf * (double)(nx - 1) / Re;
    const double tau = 0.5 + 3.0 * nu;
    return 1.0 / tau;
}

/* ------------------------------------------------------------- the step */

static int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) return nthreads;
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* Post-collision populations of every node, f*(x) = reconstruct(rho0 + stored). */
static void all_pops(int Q, int64_t n, const double *const *mom, double *F, int nt) {
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        double m[10];
        for (int k = 0; k < 10; k++) m[k] = mom[k][i];
        m[0] += 1.0;
        lbmo_reconstruct(Q, m, F + i * Q);
    }
}

static void step_with(int Q, int nx, int ny, int nz, double omega, double u_lid,
                      const double *F, double *const *mom_out, int nt) {
    double U[18] = {0};
    U[3 * 2 + 0] = u_lid;
#pragma omp parallel for collapse(2) num_threads(nt) schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double f[27], m[10];
                for (int q = 0; q < Q; q++) {
                    const int sx = x - C27[q][0], sy = y - C27[q][1], sz = z - C27[q][2];
                    const int in = sx >= 0 && sx < nx && sy >= 0 && sy < ny && sz >= 0 && sz < nz;
                    /* unknown populations are masked to zero (velocitySet.cuh:337-352) */
                    f[q] = in ? F[((int64_t)sx + (int64_t)nx * (sy + (int64_t)ny * sz)) * Q + q] : 0.0;
                }
                const int bits = lbmo_node_type(x, y, z, nx, ny, nz);
                if (bits & 0x40)
                    lbmo_moments_boundary(Q, f, bits, omega, U, m);
                else
                    lbmo_moments_interior(Q, f, m);
                lbmo_scale(m);
                lbmo_collide(m, omega);
                m[0] -= 1.0;
                const int64_t i = (int64_t)x + (int64_t)nx * (y + (int64_t)ny * z);
                for (int k = 0; k < 10; k++) mom_out[k][i] = m[k];
            }
}

int lbmo_step(int Q, int nx, int ny, int nz, double omega, double u_lid,
              const double *const *mom_in, double *const *mom_out, int nthreads) {
    if (Q != 19 && Q != 27) return -1;
    if (nx < 2 || ny < 2 || nz < 2) return -2;
    const int nt = set_threads(nthreads);
    const int64_t n = (int64_t)nx * ny * nz;
    double *F = (double *)malloc(sizeof(double) * (size_t)n * Q);
    if (!F) return -3;
    all_pops(Q, n, mom_in, F, nt);
    step_with(Q, nx, ny, nz, omega, u_lid, F, mom_out, nt);
    free(F);
    return 0;
}

int lbmo_run(int Q, int nx, int ny, int nz, double omega, double u_lid, double *const *mom,
             int nsteps, int nthreads) {
    if (Q != 19 && Q != 27) return -1;
    if (nx < 2 || ny < 2 || nz < 2) return -2;
    const int nt = set_threads(nthreads);
    const int64_t n = (int64_t)nx * ny * nz;
    double *F = (double *)malloc(sizeof(double) * (size_t)n * Q);
    if (!F) return -3;
    /* The step reads only F after all_pops, so writing mom in place is safe. */
    for (int s = 0; s < nsteps; s++) {
        all_pops(Q, n, (const double *const *)mom, F, nt);
        step_with(Q, nx, ny, nz, omega, u_lid, F, mom, nt);
    }
    free(F);
    return 0;
}

/* n steps with the function objects' running means accumulated after every
 * step, as the reference driver calls runTimeObjects.calculate(t) after the
 * step of time t (momentBasedD3Q19.cu:126-136): mean += (x - mean)/(t + 1)
 * with t = t0 + s for the s-th step of this call (timeAverage,
 * functionObjects.cuh:79-83).  x is k (kineticEnergy.cuh:71-85), the six S
 * components (strainRateTensor.cuh:71-84) and the ten stored moments
 * (moments.cuh:70-101), all from the post-step stored moments.  Any mean
 * pointer may be NULL. */
int lbmo_run_means(int Q, int nx, int ny, int nz, double omega, double u_lid, double *const *mom, int nsteps,
                   int nthreads, int64_t t0, double tau, double *kmean, double *smean, double *mmean) {
    if (Q != 19 && Q != 27) return -1;
    if (nx < 2 || ny < 2 || nz < 2) return -2;
    const int nt = set_threads(nthreads);
    const int64_t n = (int64_t)nx * ny * nz;
    double *F = (double *)malloc(sizeof(double) * (size_t)n * Q);
    if (!F) return -3;
    const double sd = 3.0 / (2.0 * 4.5 * tau), so = 3.0 / (2.0 * 9.0 * tau);
    for (int s = 0; s < nsteps; s++) {
        all_pops(Q, n, (const double *const *)mom, F, nt);
        step_with(Q, nx, ny, nz, omega, u_lid, F, mom, nt);
        const double inv = 1.0 / (double)(t0 + s + 1);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < n; i++) {
            const double u = mom[1][i], v = mom[2][i], w = mom[3][i];
            if (kmean) {
                const double k = 0.5 * sqrt(u * u + v * v + w * w);
                kmean[i] += (k - kmean[i]) * inv;
            }
            if (smean) {
                const double S[6] = {(u * u - mom[4][i]) * sd, (u * v - mom[5][i]) * so, (u * w - mom[6][i]) * so,
                                     (v * v - mom[7][i]) * sd, (v * w - mom[8][i]) * so, (w * w - mom[9][i]) * sd};
                for (int a = 0; a < 6; a++) smean[a * n + i] += (S[a] - smean[a * n + i]) * inv;
            }
            if (mmean)
                for (int a = 0; a < 10; a++) mmean[a * n + i] += (mom[a][i] - mmean[a * n + i]) * inv;
        }
    }
    free(F);
    return 0;
}

/* ------------------------------------------------------- initial fields */

void lbmo_initial_field(int nx, int ny, int nz, const double *r, double *out) {
    /* hostArray.cuh:330-350: West x==0, East x==nx-1, South y==0, North y==ny-1,
     * "front" z==0, "back" z==nz-1; value = mean over the regions a boundary node
     * lies on, internalField otherwise. */
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int w = x == 0, e = x == nx - 1, s = y == 0, n = y == ny - 1;
                const int fr = z == 0, bk = z == nz - 1;
                const int cnt = w + e + s + n + fr + bk;
                const double sum = n * r[0] + s * r[1] + e * r[2] + w * r[3] + fr * r[4] + bk * r[5];
                out[(int64_t)x + (int64_t)nx * (y + (int64_t)ny * z)] = cnt > 0 ? sum / cnt : r[6];
            }
}

/* ----------------------------------------------------- function objects */

void lbmo_k(const double *const *mom, int64_t n, double *k) {
    for (int64_t i = 0; i < n; i++) {
        const double u = mom[1][i], v = mom[2][i], w = mom[3][i];
        k[i] = sqrt(u * u + v * v + w * w) * 0.5;
    }
}

void lbmo_S(const double *const *mom, int64_t n, double tau, double *S) {
    /* S_ab = as2 (m_a m_b - m_ab) / (2 scale tau), as2 = 3, scale 4.5 (diag) or 9 */
    for (int64_t i = 0; i < n; i++) {
        const double u = mom[1][i], v = mom[2][i], w = mom[3][i];
        S[0 * n + i] = 3.0 * (u * u - mom[4][i]) / (2.0 * 4.5 * tau);
        S[1 * n + i] = 3.0 * (u * v - mom[5][i]) / (2.0 * 9.0 * tau);
        S[2 * n + i] = 3.0 * (u * w - mom[6][i]) / (2.0 * 9.0 * tau);
        S[3 * n + i] = 3.0 * (v * v - mom[7][i]) / (2.0 * 4.5 * tau);
        S[4 * n + i] = 3.0 * (v * w - mom[8][i]) / (2.0 * 9.0 * tau);
        S[5 * n + i] = 3.0 * (w * w - mom[9][i]) / (2.0 * 4.5 * tau);
    }
}

void lbmo_time_average(double *mean, const double *x, int64_t n, int64_t t) {
    const double inv = 1.0 / (double)(t + 1);
    for (int64_t i = 0; i < n; i++) mean[i] = mean[i] + (x[i] - mean[i]) * inv;
}

/* ------------------------------------------- reference tile/face emulation */

/* Load table of halo::load (halo.cuh:208-291): for each tile side, the
 * populations read from the neighbour's opposite face, with slot and the
 * shift (-1 = "m1", +1 = "p1") of the two in-face coordinates. */
typedef struct {
    int q, slot, du, dv;
} load_t;

static const load_t LOAD_X0[9] = {{1, 0, 0, 0},   {7, 1, -1, 0},  {9, 2, 0, -1},
                                  {13, 3, 1, 0},  {15, 4, 0, 1},  {19, 5, -1, -1},
                                  {21, 6, -1, 1}, {23, 7, 1, -1}, {26, 8, 1, 1}};
static const load_t LOAD_X1[9] = {{2, 0, 0, 0},   {8, 1, 1, 0},   {10, 2, 0, 1},
                                  {14, 3, -1, 0}, {16, 4, 0, -1}, {20, 5, 1, 1},
                                  {22, 6, 1, -1}, {24, 7, -1, 1}, {25, 8, -1, -1}};
static const load_t LOAD_Y0[9] = {{3, 0, 0, 0},   {7, 1, -1, 0},  {11, 2, 0, -1},
                                  {14, 3, 1, 0},  {17, 4, 0, 1},  {19, 5, -1, -1},
                                  {21, 6, -1, 1}, {24, 7, 1, 1},  {25, 8, 1, -1}};
static const load_t LOAD_Y1[9] = {{4, 0, 0, 0},   {8, 1, 1, 0},   {12, 2, 0, 1},
                                  {13, 3, -1, 0}, {18, 4, 0, -1}, {20, 5, 1, 1},
                                  {22, 6, 1, -1}, {23, 7, -1, -1}, {26, 8, -1, 1}};
static const load_t LOAD_Z0[9] = {{5, 0, 0, 0},   {9, 1, -1, 0},  {11, 2, 0, -1},
                                  {16, 3, 1, 0},  {18, 4, 0, 1},  {19, 5, -1, -1},
                                  {22, 6, 1, 1},  {23, 7, -1, 1}, {25, 8, 1, -1}};
static const load_t LOAD_Z1[9] = {{6, 0, 0, 0},   {10, 1, 1, 0},  {12, 2, 0, 1},
                                  {15, 3, -1, 0}, {17, 4, 0, -1}, {20, 5, 1, 1},
                                  {21, 6, -1, -1}, {24, 7, 1, -1}, {26, 8, -1, 1}};

static int wrap(int a, int n) { return ((a % n) + n) % n; }

int lbmo_tile_source_map(int Q, int nx, int ny, int nz, int64_t *src, uint8_t *known) {
    const int qf = lbmo_qf(Q);
    if (qf < 0) return -1;
    if (nx % 8 || ny % 8 || nz % 8) return -2;
    const int T = 8;
    const int NBX = nx / T, NBY = ny / T, NBZ = nz / T;
    const int64_t nfaces = (int64_t)NBX * NBY * NBZ * T * T * qf;
    /* save maps: face buffer address -> (global node, population), filled by the
     * face writes of every tile (halo.cuh:804-905 layout, all six faces). */
    int64_t *saved_node[6];
    int *saved_pop[6];
    for (int f = 0; f < 6; f++) {
        saved_node[f] = (int64_t *)malloc(sizeof(int64_t) * nfaces);
        saved_pop[f] = (int *)malloc(sizeof(int) * nfaces);
        if (!saved_node[f] || !saved_pop[f]) return -3;
        for (int64_t i = 0; i < nfaces; i++) {
            saved_node[f][i] = -1;
            saved_pop[f][i] = -1;
        }
    }
#define TILE(bx, by, bz) ((int64_t)(bx) + (int64_t)NBX * ((by) + (int64_t)NBY * (bz)))
#define IDXPOP(a, b, s, tile) ((a) + T * ((b) + T * ((s) + (int64_t)qf * (tile))))
#define GID(x, y, z) ((int64_t)(x) + (int64_t)nx * ((y) + (int64_t)ny * (z)))
    for (int bz = 0; bz < NBZ; bz++)
        for (int by = 0; by < NBY; by++)
            for (int bx = 0; bx < NBX; bx++)
                for (int a = 0; a < T; a++)
                    for (int b = 0; b < T; b++)
                        for (int s = 0; s < qf; s++) {
                            const int64_t t = TILE(bx, by, bz);
                            /* x faces: (ty,tz) = (a,b); y faces: (tx,tz); z faces: (tx,ty) */
                            const int64_t ad = IDXPOP(a, b, s, t);
                            saved_node[0][ad] = GID(bx * T + 0, by * T + a, bz * T + b);
                            saved_node[1][ad] = GID(bx * T + T - 1, by * T + a, bz * T + b);
                            saved_node[2][ad] = GID(bx * T + a, by * T + 0, bz * T + b);
                            saved_node[3][ad] = GID(bx * T + a, by * T + T - 1, bz * T + b);
                            saved_node[4][ad] = GID(bx * T + a, by * T + b, bz * T + 0);
                            saved_node[5][ad] = GID(bx * T + a, by * T + b, bz * T + T - 1);
                            for (int f = 0; f < 6; f++) saved_pop[f][ad] = SLOTS[f][s];
                        }
    int err = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int tx = x % T, ty = y % T, tz = z % T;
                const int bx = x / T, by = y / T, bz = z / T;
                const int64_t d = GID(x, y, z);
                const int bits = lbmo_node_type(x, y, z, nx, ny, nz);
                int64_t s_of[27];
                /* in-tile pull with wrap inside the tile (streaming.cuh:126-154) */
                for (int q = 0; q < Q; q++) {
                    const int sx = wrap(tx - C27[q][0], T), sy = wrap(ty - C27[q][1], T),
                              sz = wrap(tz - C27[q][2], T);
                    s_of[q] = GID(bx * T + sx, by * T + sy, bz * T + sz);
                }
                /* halo loads, x then y then z; block indices wrap periodically */
                const int bxm1 = wrap(bx - 1, NBX), bxp1 = wrap(bx + 1, NBX);
                const int bym1 = wrap(by - 1, NBY), byp1 = wrap(by + 1, NBY);
                const int bzm1 = wrap(bz - 1, NBZ), bzp1 = wrap(bz + 1, NBZ);
                struct side {
                    int active, face;
                    const load_t *tab;
                } sides[6] = {{tx == 0, 1, LOAD_X0},     {tx == T - 1, 0, LOAD_X1},
                              {ty == 0, 3, LOAD_Y0},     {ty == T - 1, 2, LOAD_Y1},
                              {tz == 0, 5, LOAD_Z0},     {tz == T - 1, 4, LOAD_Z1}};
                /* reference applies x0 | x1 (else-if), then y0 | y1, then z0 | z1 */
                for (int k = 0; k < 6; k++) {
                    if (!sides[k].active) continue;
                    if ((k & 1) && sides[k - 1].active) continue; /* else-if branch */
                    const int axis = k / 2;
                    for (int j = 0; j < qf; j++) {
                        const load_t L = sides[k].tab[j];
                        int ua, va, ub[3] = {bx, by, bz};
                        /* in-face coordinates and their tile adjustment */
                        int u0, v0, uax, vax;
                        if (axis == 0) { u0 = ty; v0 = tz; uax = 1; vax = 2; }
                        else if (axis == 1) { u0 = tx; v0 = tz; uax = 0; vax = 2; }
                        else { u0 = tx; v0 = ty; uax = 0; vax = 1; }
                        ua = wrap(u0 + L.du, T);
                        va = wrap(v0 + L.dv, T);
                        const int nb[3] = {NBX, NBY, NBZ};
                        const int mb[3] = {bxm1, bym1, bzm1}, pb[3] = {bxp1, byp1, bzp1};
                        (void)nb;
                        if (L.du < 0 && u0 == 0) ub[uax] = mb[uax];
                        if (L.du > 0 && u0 == T - 1) ub[uax] = pb[uax];
                        if (L.dv < 0 && v0 == 0) ub[vax] = mb[vax];
                        if (L.dv > 0 && v0 == T - 1) ub[vax] = pb[vax];
                        ub[axis] = (k & 1) ? pb[axis] : mb[axis];
                        const int64_t ad = IDXPOP(ua, va, L.slot, TILE(ub[0], ub[1], ub[2]));
                        if (saved_pop[sides[k].face][ad] != L.q) err = -4;
                        s_of[L.q] = saved_node[sides[k].face][ad];
                    }
                }
                for (int q = 0; q < Q; q++) {
                    src[d * Q + q] = s_of[q];
                    known[d * Q + q] = (uint8_t)lbmo_known(Q, q, bits);
                }
            }
    for (int f = 0; f < 6; f++) {
        free(saved_node[f]);
        free(saved_pop[f]);
    }
#undef TILE
#undef IDXPOP
#undef GID
    return err;
}
