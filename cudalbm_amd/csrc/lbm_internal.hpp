// lbm_internal.hpp -- launchers and geometry helpers shared by the C-ABI units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "step_kernel.hpp"

namespace lbm {

bool tile_supported(int tx, int ty);
void launch_step(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool fo,
                 bool probe);
void launch_fill(int Q, int tx, int ty, const StepParams &p, int grid, hipStream_t s, int storage, bool probe);
// Words of one face-buffer set and the offsets of its Y and Z blocks.
int64_t face_words(int Q, int nx, int ny, int nz, int tx, int ty, int lz, int64_t *yoff, int64_t *zoff);
int choose_lz(int nx, int ny, int nz, int tx, int ty);
void choose_tile(int Q, int nx, int ny, int *tx, int *ty);
std::vector<int64_t> make_perm(int layout, int nx, int ny, int nz, int tx, int ty);

}  // namespace lbm
