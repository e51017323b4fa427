/*
 * lbm_oracle.h -- FP64 scalar CPU restatement of the reference moment-representation
 * LBM step (D3Q19 / D3Q27, second-order regularised collision, lid-driven cavity).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the HIP
 * solver and the CPU baseline that bench.py reports beside it.  Nothing in the
 * product path (cudalbm_amd, include/lbm_amd.h) links, loads or calls it.
 *
 * Conventions shared with the reference (cited per function in lbm_oracle.c):
 *   - 10 stored moments per node, in this order:
 *       0 rho-1, 1 3u, 2 3v, 3 3w, 4 4.5(Pxx/rho-cs2), 5 9Pxy/rho, 6 9Pxz/rho,
 *       7 4.5(Pyy/rho-cs2), 8 9Pyz/rho, 9 4.5(Pzz/rho-cs2)
 *     (globalFunctions.cuh:482-503, velocitySet.cuh:130-142)
 *   - node-type bitmask: bit0 x==0 West, bit1 x==nx-1 East, bit2 y==0 South,
 *     bit3 y==ny-1 North, bit4 z==0 Back, bit5 z==nz-1 Front, bit6 any
 *     (normalVector.cuh:300-334)
 *   - arrays handed to the oracle are in global x-fastest order,
 *     idx = x + nx*(y + ny*z).
 *
 * Parity status per function (how each is pinned) is recorded in the header of
 * lbm_oracle.c and in the tests under tests/.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Velocity-set tables. q in [0,Q). Returns 0 on success, -1 on a bad Q. */
int lbmo_velocity_set(int Q, int *cx, int *cy, int *cz, double *w);

/* Number of populations per face slot set (5 for D3Q19, 9 for D3Q27). */
int lbmo_qf(int Q);

/* Reference face-slot lists: for face f in {x0,x1,y0,y1,z0,z1} the QF
 * population indices stored in that face buffer, in slot order
 * (haloFace.cuh:329-448, halo.cuh:813-904).  out has 6*QF entries. */
int lbmo_face_slots(int Q, int *out);

/* f_q = rho w_q [pics2 + ...] from (full rho, 9 scaled moments) (D3Q19.cuh:346-374,
 * D3Q27.cuh:294-331).  m[0] is the FULL density here. */
int lbmo_reconstruct(int Q, const double *m, double *f);

/* Unmasked interior moments (unscaled; m[0] = full rho) (velocitySet.cuh:398-417). */
int lbmo_moments_interior(int Q, const double *f, double *m);

/* incomingSwitch: 1 if population q is known at a node of type `bits`
 * (velocitySet.cuh:204-220). */
int lbmo_known(int Q, int q, int bits);

/* Boundary moments (unscaled, m[0] = full rho) for a node of type `bits`
 * (boundaryConditions.cuh:99-554).  U[6][3] are the face velocities in the
 * reference order West, East, North, South, Front, Back. */
int lbmo_moments_boundary(int Q, const double *f, int bits, double omega,
                          const double *U, double *m);

/* Scale (velocitySet.cuh:130-142) and collide (secondOrder.cuh:93-105), in place. */
void lbmo_scale(double *m);
void lbmo_collide(double *m, double omega);

/* Node-type bitmask (normalVector.cuh:321-334). */
int lbmo_node_type(int x, int y, int z, int nx, int ny, int nz);

/* omega from the case parameters (latticeMesh.cuh:178-182). */
double lbmo_omega(double u_inf, double Re, int nx);

/* One full step over the domain: mom_in (stored moments, 10 arrays of
 * nx*ny*nz) -> mom_out.  u_lid is U_North[0]; every other face is at rest
 * (latticeMesh.cuh:183-188).  nthreads <= 0 uses all OpenMP threads.
 * Returns 0 on success. */
int lbmo_step(int Q, int nx, int ny, int nz, double omega, double u_lid,
              const double *const *mom_in, double *const *mom_out, int nthreads);

/* n steps, ping-ponging internally; result lands in mom (in place). */
int lbmo_run(int Q, int nx, int ny, int nz, double omega, double u_lid,
             double *const *mom, int nsteps, int nthreads);

/* lbmo_run with the function objects' FP64 running means updated after every
 * step: kmean (n), smean (6*n, xx,xy,xz,yy,yz,zz) and mmean (10*n, the stored
 * moments), each may be NULL; t0 = absolute index of the first step
 * (mean += (x - mean)/(t0 + s + 1)); tau as in lbmo_S. */
int lbmo_run_means(int Q, int nx, int ny, int nz, double omega, double u_lid,
                   double *const *mom, int nsteps, int nthreads, int64_t t0, double tau,
                   double *kmean, double *smean, double *mmean);

/* Initial field for one moment from its 7 region values (already scaled as the
 * reference scales them): order North, South, East, West, Front, Back,
 * internalField (hostArray.cuh:304-359).  out is x-fastest global order. */
void lbmo_initial_field(int nx, int ny, int nz, const double *region7, double *out);

/* Function objects on stored (scaled) moments (kineticEnergy.cuh:71-85,
 * strainRateTensor.cuh:71-84, functionObjects.cuh:79-83).
 * k has n entries; S has 6*n entries in the order xx,xy,xz,yy,yz,zz. */
void lbmo_k(const double *const *mom, int64_t n, double *k);
void lbmo_S(const double *const *mom, int64_t n, double tau, double *S);
void lbmo_time_average(double *mean, const double *x, int64_t n, int64_t t);

/* Emulation of the reference 8^3-tile streaming + face-buffer load index
 * logic (streaming.cuh:126-154, halo.cuh:171-293, globalFunctions.cuh:415-476).
 * For every destination node d (global x-fastest index) and population q,
 * src[d*Q+q] = global index of the node whose post-collision population q is
 * pulled, and known[d*Q+q] = incomingSwitch at d.  nx,ny,nz multiples of 8. */
int lbmo_tile_source_map(int Q, int nx, int ny, int nz, int64_t *src, uint8_t *known);

#ifdef __cplusplus
}
#endif

#endif
