// Device view of the tree for the FP64 upward pass (pgrav_moments.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgrav {

struct MomentsView {
    const double4 *bpos = nullptr;  // sorted bodies (x, y, z, m)
    const int64_t *first = nullptr;  // per node: first body
    const int *count = nullptr;      // bodies below
    const int *child = nullptr;      // first daughter (-1: leaf); daughters are contiguous
    const int *nchild = nullptr;
    const int *lev_nodes = nullptr;  // node indices grouped by level (lev_off on the host)
    double4 *cpos = nullptr;         // out: centre of mass and mass
    double *bmax = nullptr;          // out
    double *raw = nullptr;           // out: 35 raw moments about the centre per node
    float *mom32 = nullptr;          // out: prepared records
    double *mom64 = nullptr;         // out (may be null)
};

hipError_t launch_moments(const MomentsView &T, const int *lev_off, int nlev, int64_t nnodes,
                          int order, hipStream_t st);

}  // namespace pgrav
