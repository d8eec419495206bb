// Bodies in and results out of a plan, on the device: the caller's order <->
// the plan's sorted slots (slot i holds caller body perm[i]).
//
// k_gather_bodies   caller-order FP64 positions and masses -> sorted double4
//                   slots.  Periodic plans keep every body at the image nearest
//                   its current slot position (the tree, groups and lists are
//                   fixed, so a body that crosses the box boundary must stay
//                   beside the cell it belongs to; the kernels take the
//                   nearest image of every difference).  The smallest caller
//                   index of a body with a non-positive mass or a non-finite
//                   value is reported through *bad.
// k_scatter4        sorted float4 / double4 records -> caller order
// k_gather4         caller-order float / double quadruples -> sorted float4
//
// HBM-bound, one thread per body.
#include <math.h>

#include "pgrav_batch.h"

namespace pgrav {

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

__global__ __launch_bounds__(256) void k_gather_bodies(const double *__restrict__ pos3,
                                                       const double *__restrict__ mass,
                                                       const int64_t *__restrict__ perm, int64_t n,
                                                       int periodic, double box,
                                                       const double4 *__restrict__ prev,
                                                       double4 *__restrict__ out,
                                                       unsigned long long *bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int64_t j = perm[i];
    double x = pos3[3 * j], y = pos3[3 * j + 1], z = pos3[3 * j + 2];
    const double m = mass[j];
    if (!(m > 0.0) || !isfinite(m) || !isfinite(x) || !isfinite(y) || !isfinite(z)) {
        atomicMin(bad, (unsigned long long)j);
        return;
    }
    if (periodic) {
        const double4 p = prev[i];
        x -= box * rint((x - p.x) / box);
        y -= box * rint((y - p.y) / box);
        z -= box * rint((z - p.z) / box);
    }
    out[i] = make_double4(x, y, z, m);
}

template <typename S, typename D>
__global__ __launch_bounds__(256) void k_scatter4(const S *__restrict__ src,
                                                  const int64_t *__restrict__ perm, int64_t n,
                                                  D *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const S s = src[i];
    D d;
    d.x = s.x;
    d.y = s.y;
    d.z = s.z;
    d.w = s.w;
    dst[perm[i]] = d;
}

template <typename S>
__global__ __launch_bounds__(256) void k_gather4(const S *__restrict__ src,
                                                 const int64_t *__restrict__ perm, int64_t n,
                                                 float4 *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const S s = src[perm[i]];
    dst[i] = make_float4((float)s.x, (float)s.y, (float)s.z, (float)s.w);
}

static dim3 grid_of(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_gather_bodies(const double *pos3, const double *mass, const int64_t *perm,
                                int64_t n, int periodic, double box, const double4 *prev,
                                double4 *out, unsigned long long *bad, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, bad, ~0ull);
    if (n > 0)
        hipLaunchKernelGGL(k_gather_bodies, grid_of(n), dim3(256), 0, st, pos3, mass, perm, n,
                           periodic, box, prev, out, bad);
    return hipGetLastError();
}

hipError_t launch_scatter4(const void *src, int src_fp64, const int64_t *perm, int64_t n, void *dst,
                           int dst_fp64, hipStream_t st) {
    if (n <= 0)
        return hipSuccess;
    if (src_fp64 && dst_fp64)
        hipLaunchKernelGGL((k_scatter4<double4, double4>), grid_of(n), dim3(256), 0, st,
                           (const double4 *)src, perm, n, (double4 *)dst);
    else if (src_fp64)
        hipLaunchKernelGGL((k_scatter4<double4, float4>), grid_of(n), dim3(256), 0, st,
                           (const double4 *)src, perm, n, (float4 *)dst);
    else if (dst_fp64)
        hipLaunchKernelGGL((k_scatter4<float4, double4>), grid_of(n), dim3(256), 0, st,
                           (const float4 *)src, perm, n, (double4 *)dst);
    else
        hipLaunchKernelGGL((k_scatter4<float4, float4>), grid_of(n), dim3(256), 0, st,
                           (const float4 *)src, perm, n, (float4 *)dst);
    return hipGetLastError();
}

hipError_t launch_gather4(const void *src, int src_fp64, const int64_t *perm, int64_t n,
                          float4 *dst, hipStream_t st) {
    if (n <= 0)
        return hipSuccess;
    if (src_fp64)
        hipLaunchKernelGGL(k_gather4<double4>, grid_of(n), dim3(256), 0, st, (const double4 *)src,
                           perm, n, dst);
    else
        hipLaunchKernelGGL(k_gather4<float4>, grid_of(n), dim3(256), 0, st, (const float4 *)src,
                           perm, n, dst);
    return hipGetLastError();
}

}  // namespace pgrav
