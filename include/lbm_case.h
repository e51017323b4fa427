/*
 * lbm_case.h -- host-only C ABI for the reference's case interface and output
 * format: case files (programControl, latticeMesh, initialConditions,
 * functionObjects) and .LBMBin files.  No GPU is touched by these calls.
 *
 * Every function returns 0 on success or a negative code; lbm_case_last_error()
 * gives the message (thread-local).  The reference throws std::runtime_error or
 * exits instead (src/strings.cuh:421-453, src/errorHandler.cuh:142-168).
 */
#ifndef LBM_CASE_H
#define LBM_CASE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_case_info {
    char case_name[256];
    double Re, u_inf;
    int64_t n_time_steps, save_interval, info_interval;
    int nx, ny, nz;
    double Lx, Ly, Lz;
    int64_t latest_time; /* highest <caseName>_<t>.LBMBin in the case directory, 0 if none */
} lbm_case_info;

const char *lbm_case_last_error(void);

/* Replaces programControl(argc, argv) case-file reads (src/programControl/programControl.cuh:69-77),
 * host::latticeMesh's file reads (src/latticeMesh/latticeMesh.cuh:85-93) and
 * fileIO::latestTime (src/fileIO/fileIO.cuh:222-233), relative to `dir`. */
int lbm_case_read(const char *dir, lbm_case_info *out);

/* Value of `name` by the reference rule: first line containing the substring,
 * split on spaces, second token without its last character
 * (src/strings.cuh:561-595).  `out` receives at most cap-1 characters. */
int lbm_case_parameter(const char *file, const char *name, char *out, int cap);

/* Initial stored moments from <dir>/initialConditions (src/array/hostArray.cuh:304-359,
 * src/boundaryConditions/boundaryValue.cuh:147-254) into 10 host arrays in
 * global x-fastest order (idx = x + nx*(y + ny*z)). */
int lbm_case_initial_moments(const char *dir, int nx, int ny, int nz, float *const m[10]);

/* Enabled function objects from <dir>/functionObjects (src/functionObjects/functionObjects.cuh:104-107):
 * bit 0 k, 1 kMean, 2 S, 3 SMean (the lbm_amd.h LBM_FO_* bits), 4 momentsMean. */
int lbm_case_function_objects(const char *dir, unsigned *mask);

/* Writes a .LBMBin file (src/fileIO/fileOutput.cuh:82-183): ASCII header then
 * the nvars fields as FP32, var-major, each in the reference's 8^3-tiled order
 * (src/memory/memory.cuh:157-195).  Fields are given in global x-fastest
 * order; time_average selects timeType "timeAverage". */
int lbm_lbmbin_write(const char *path, int nx, int ny, int nz, int nvars, const char *const *names,
                     const float *const *fields, int64_t time_step, int time_average);

/* Header of a .LBMBin file (src/fileIO/fileInput.cuh:99-489). names receives
 * nvars NUL-separated names in a buffer of names_cap bytes (may be NULL). */
int lbm_lbmbin_header(const char *path, int *nx, int *ny, int *nz, int *nvars, int *scalar_bytes,
                      int64_t *time_step, char *names, int names_cap);

/* One named field of a .LBMBin file, untiled to global x-fastest order
 * (src/fileIO/fileInput.cuh:608-822). */
int lbm_lbmbin_read_field(const char *path, const char *name, float *out);

#ifdef __cplusplus
}
#endif

#endif
