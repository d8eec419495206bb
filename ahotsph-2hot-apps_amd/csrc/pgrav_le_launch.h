// Launcher of the far-field pass of periodic = 2 plans (pgrav_ewald_le.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgrav {

#define LE_MAX_BLOCKS 4096  // workgroups of the moment pass (partial sums: 165 doubles each)

// Adds the far images and background of the explicit-image scheme, times G, to
// the result of the sorted bodies bpos[0..n) (exactly one of out32 / out64).
// Wt: the LE_NMON x LE_NMON matrix of le_matrix, transposed; partial: (nblocks + 1) x
// LE_NMON (the last row receives the global moments); coef: LE_NMON.
hipError_t launch_ewald_le(const double4 *bpos, int64_t n, double box, double G, const double *Wt,
                           double *partial, int nblocks, double *coef, float4 *out32,
                           double4 *out64, hipStream_t st);

}  // namespace pgrav
