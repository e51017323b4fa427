// lbm_internal.hpp -- launchers and geometry helpers shared by the C-ABI units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "step_kernel.hpp"

namespace lbm {

// Failures inside the library are C++ exceptions, turned into status codes at
// the C ABI (no exception crosses it).
struct HipError {
    hipError_t e;
    const char *what;
};
struct ArgError {
    int code;
    std::string msg;
};
#define HIPX(expr)                                       \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) throw ::lbm::HipError{e_, #expr}; \
    } while (0)

// Relaxation of a cavity case, latticeMesh.cuh:178-182 evaluated in FP64:
// nu = u (nx-1)/Re, tau = 0.5 + 3 nu, omega = 1/tau.
inline void relaxation(double Re, double u_inf, int nx, double *omega, double *tau) {
    const double nu = u_inf * (double)(nx - 1) / Re;
    *tau = 0.5 + 3.0 * nu;
    *omega = 1.0 / *tau;
}

// Footprints the single-case step is built for (D3Q19 also 32x32 and 64x16).
bool tile_supported(int Q, int tx, int ty);
void launch_step(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool fo,
                 bool probe);
void launch_fill(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool probe);
// Words of one face-buffer set and the offsets of its Y and Z blocks.
int64_t face_words(int Q, int nx, int ny, int nz, int tx, int ty, int lz, int64_t *yoff, int64_t *zoff);
int choose_lz(int Q, int nx, int ny, int nz, int tx, int ty, int64_t slots, bool fp16, int mean_bytes);
int64_t resident_workgroups(int Q, int tx, int ty, int device, int storage, bool fo);
void choose_tile(int Q, int nx, int ny, bool fp16, bool means, int *tx, int *ty);
std::vector<int64_t> make_perm(int layout, int nx, int ny, int nz, int tx, int ty);

}  // namespace lbm

struct lbm_handle;
namespace lbm {
// Mesh size of a solver handle (library-internal accessor).
void handle_dims(const lbm_handle *h, int *nx, int *ny, int *nz);
// Warm start on the device: the equilibrium moments of a coarse S^3 field
// (device memory, 4 channels u/U, v/U, w/U, p'/U^2 at stride cstride floats per
// point) interpolated onto the handle's nodes, then its face buffers rebuilt.
void warm_moments(lbm_handle *h, const float *field, int S, int cstride, double U);

}  // namespace lbm
