/*
 * lbm_post.h -- host-only post-processing of .LBMBin results: the CPU
 * counterparts of the reference's fieldConvert and fieldCalculate
 * (applications/fieldConvert, applications/fieldCalculate), which need no GPU
 * here.  Fields are FP32 arrays in global x-fastest order (idx = x + nx*(y + ny*z)),
 * as lbm_lbmbin_read_field returns them.
 *
 * Every function returns 0 (or a count) on success and a negative lbm_status
 * on error; lbm_case_last_error() (lbm_case.h) gives the message.
 */
#ifndef LBM_POST_H
#define LBM_POST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Writer kinds of postProcess::writers (src/postProcess/writerFunction.cuh:70-73). */
typedef enum { LBM_POST_VTU = 0, LBM_POST_VTS = 1, LBM_POST_TECPLOT = 2 } lbm_post_format;

/* Sorted time indices t of the files <prefix>_<t>.LBMBin in `dir`
 * (fileIO::timeIndices, src/fileIO/fileIO.cuh:98-221).  Writes at most cap
 * values; returns the number of files found. */
int lbm_post_time_indices(const char *dir, const char *prefix, int64_t *out, int cap);

/* Writes `path` + extension (.vtu, .vts, .dat) with the node coordinates
 * x = Lx i/(nx-1) (meshCoordinates, src/postProcess/postProcess.cuh:74-99) and
 * the nvars point fields: VTK XML with raw appended blocks of UInt64 byte
 * counts (VTU.cuh:65-136, VTS.cuh:65-116) or Tecplot ASCII FEBRICK blocks
 * (Tecplot.cuh:65-118). */
int lbm_post_write(int format, const char *path, int nx, int ny, int nz, double Lx, double Ly, double Lz, int nvars,
                   const char *const *names, const float *const *fields);

/* Difference schemes for the derivatives below.  LBM_POST_CENTRAL is the
 * central difference of the given order (coefficients of the differences
 * f(x+k) - f(x-k): order 4: 2/3, -1/12; 6: 3/4, -3/20, 1/60; 8: 4/5, -1/5,
 * 4/105, -1/280).  LBM_POST_REFERENCE reproduces the reference's stencils
 * (src/numericalSchemes/derivativeSchemes/derivativeSchemes.cuh:107-139), which
 * add every term with a plus sign: identical to central at order 2, but at
 * orders 4-8 they are not consistent (d/dx of f = x gives 5/3, 2.2, 2.657),
 * so its output is the reference's output, not a derivative. */
typedef enum { LBM_POST_CENTRAL = 0, LBM_POST_REFERENCE = 1 } lbm_post_scheme;

/* d f / d(axis) with unit spacing, order 2, 4, 6 or 8, with the reference's
 * odd extension about the end nodes: f(-j) = -f(j), f(n-1+j) = -f(n-1-j)
 * (derivativeSchemes.cuh:68-106).  axis 0 = x, 1 = y, 2 = z.  FP64 out. */
int lbm_post_derivative(int scheme, int order, int axis, int nx, int ny, int nz, const float *f, double *out);

/* Curl (3 arrays) and divergence of (u, v, w) with the scheme above, FP32 out
 * (curl.cuh:57-97, div.cuh:57-89). */
int lbm_post_curl(int scheme, int order, int nx, int ny, int nz, const float *u, const float *v, const float *w,
                  float *cx, float *cy, float *cz);
int lbm_post_div(int scheme, int order, int nx, int ny, int nz, const float *u, const float *v, const float *w,
                 float *out);

/* Cut plane at coordinate `c` normal to `axis`, following extractCutPlane
 * (applications/fieldConvert/fieldConvert.cuh:262-330) exactly: index
 * i = (n-1) (c L), i0 = floor(i), i1 = ceil(i), weight = c - i0, value =
 * (1 - weight) f(i0) + weight f(i1); the plane is written in the order of the
 * two remaining axes (first index fastest).  out has the plane's node count. */
int lbm_post_cut_plane(int axis, double c, int nx, int ny, int nz, double L, const float *f, float *out);

#ifdef __cplusplus
}
#endif

#endif
