// Far field of the explicit-image periodic scheme on the GPU, in FP64: the
// pass nln/main.c:1811-1862 (FixCubeEwaldLE) runs on the host after the walk.
// Four launches, all graph-capturable (no host round trip, no memset):
//
//   k_le_partials  global moments mu_gamma = sum_j m_j x_j^gamma (165 monomials,
//                  degree <= 8, box units about the box centre): each workgroup
//                  sums a contiguous range of the sorted bodies, 128 at a time
//                  staged in LDS as m x^a, y^b, z^c; one thread per monomial
//   k_le_reduce    fixed-order sum of the workgroups' partials, one workgroup
//                  per monomial
//   k_le_coeffs    C = W mu (W: the lattice expansion and background, built on
//                  the host)
//   k_le_apply     per body: the degree-8 polynomial sum C_beta x^beta and its
//                  gradient, added to the force result with the G / L and
//                  G / L^2 scales and the self term -xi m_i (nln/ewald_le2.c:44)
//
// Algorithmic work: 165 x 2 flops per body for the moments, ~4 x 165 FMA per
// body for the evaluation; 32 B read per body per pass plus the result
// read-modify-write -- a few tenths of a millisecond at N = 16.8 M.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <vector>

#include "../../include/pgrav.h"
#include "pgrav_host.h"
#include "pgrav_internal.h"
#include "pgrav_le.h"
#include "pgrav_le_launch.h"

namespace pgrav {

namespace {

// The bodies as the kernels read them: sorted double4 (x, y, z, m) of a plan, or
// the caller's separate FP64 positions (n x 3) and masses.
struct Bodies4 {
    const double4 *__restrict__ b;
    __device__ double4 operator()(int64_t i) const { return b[i]; }
};
struct Bodies3 {
    const double *__restrict__ p, *__restrict__ m;
    __device__ double4 operator()(int64_t i) const {
        return make_double4(p[3 * i], p[3 * i + 1], p[3 * i + 2], m[i]);
    }
};

constexpr int LE_TILE = 128;
template <typename B>
__global__ __launch_bounds__(256) void k_le_partials(B bpos, int64_t n, double box, int64_t chunk,
                                                     double *__restrict__ partial) {
    __shared__ double pm[LE_TILE][LE_DEG + 1], py[LE_TILE][LE_DEG + 1], pz[LE_TILE][LE_DEG + 1];
    const int t = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    int ea = 0, eb = 0, ec = 0;
    if (t < LE_NMON) {
        // the thread's monomial (a runtime index into the LDS tables)
#pragma unroll
        for (int k = 0; k < LE_NMON; k++)
            if (k == t) {
                ea = LE_MON.e[k][0];
                eb = LE_MON.e[k][1];
                ec = LE_MON.e[k][2];
            }
    }
    const double ibox = 1.0 / box;
    double acc = 0.0;
    for (int64_t s = b0; s < b1; s += LE_TILE) {
        const int m = (int)min((int64_t)LE_TILE, b1 - s);
        if (t < m) {
            const double4 b = bpos(s + t);
            const double x = (b.x - 0.5 * box) * ibox, y = (b.y - 0.5 * box) * ibox,
                         z = (b.z - 0.5 * box) * ibox;
            double px = b.w, qy = 1.0, qz = 1.0;
            for (int k = 0; k <= LE_DEG; k++) {
                pm[t][k] = px;
                py[t][k] = qy;
                pz[t][k] = qz;
                px *= x;
                qy *= y;
                qz *= z;
            }
        }
        __syncthreads();
        if (t < LE_NMON)
            for (int j = 0; j < m; j++) acc += pm[j][ea] * py[j][eb] * pz[j][ec];
        __syncthreads();
    }
    if (t < LE_NMON)
        partial[(size_t)blockIdx.x * LE_NMON + t] = acc;
}

// mu_k = sum over the workgroups' partials, one workgroup per monomial (fixed order)
__global__ __launch_bounds__(256) void k_le_reduce(const double *__restrict__ partial, int nblocks,
                                                   double *__restrict__ mu) {
    __shared__ double r[256];
    const int k = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int b = t; b < nblocks; b += blockDim.x) s += partial[(size_t)b * LE_NMON + k];
    r[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            r[t] += r[t + h];
        __syncthreads();
    }
    if (t == 0)
        mu[k] = r[0];
}

// C = W mu; Wt is W transposed (Wt[g][beta]) so that the threads read along a row
// sb: coef[LE_NMON + a] = (4 pi / 3) times the dipole moment mu_{e_a}: the
// acceleration-only term of the subtract_background form (nln/ewald_le2.c:371-375)
__global__ __launch_bounds__(256) void k_le_coeffs(const double *__restrict__ mu,
                                                   const double *__restrict__ Wt,
                                                   double *__restrict__ coef, int sb) {
    __shared__ double m[LE_NMON];
    const int t = threadIdx.x;
    if (t < LE_NMON)
        m[t] = mu[t];
    __syncthreads();
    if (t < LE_NMON) {
        double s = 0.0;
        for (int g = 0; g < LE_NMON; g++) s += Wt[(size_t)g * LE_NMON + t] * m[g];
        coef[t] = s;
    }
    if (sb && t < 3)
        coef[LE_NMON + t] = 4.0 * M_PI / 3.0 * m[le_index(t == 0, t == 1, t == 2)];
}

template <typename B, typename V>
__global__ __launch_bounds__(256) void k_le_apply(B bpos, int64_t n, double box, double G,
                                                  const double *__restrict__ coef, int sb,
                                                  V *__restrict__ out) {
    __shared__ double c[LE_NMON];
    for (int k = threadIdx.x; k < LE_NMON; k += blockDim.x) c[k] = coef[k];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double4 b = bpos(i);
    const double ibox = 1.0 / box;
    double px[LE_DEG + 1], py[LE_DEG + 1], pz[LE_DEG + 1];
    px[0] = py[0] = pz[0] = 1.0;
    const double x = (b.x - 0.5 * box) * ibox, y = (b.y - 0.5 * box) * ibox, z = (b.z - 0.5 * box) * ibox;
#pragma unroll
    for (int k = 1; k <= LE_DEG; k++) {
        px[k] = px[k - 1] * x;
        py[k] = py[k - 1] * y;
        pz[k] = pz[k - 1] * z;
    }
    double phi = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int k = 0; k < LE_NMON; k++) {
        const int a = LE_MON.e[k][0], bb = LE_MON.e[k][1], cc = LE_MON.e[k][2];
        const double yz = py[bb] * pz[cc];
        phi += c[k] * px[a] * yz;
        if (a)
            gx += c[k] * a * px[a - 1] * yz;
        if (bb)
            gy += c[k] * bb * px[a] * py[bb - 1] * pz[cc];
        if (cc)
            gz += c[k] * cc * px[a] * py[bb] * pz[cc - 1];
    }
    if (sb) {
        // the dipole term on the acceleration only; no self term (it belongs to
        // ewald_background, which the subtract_background form does not call)
        gx += coef[LE_NMON];
        gy += coef[LE_NMON + 1];
        gz += coef[LE_NMON + 2];
    }
    const double sa = -G * ibox * ibox, sp = G * ibox;
    V r = out[i];
    r.x += sa * gx;
    r.y += sa * gy;
    r.z += sa * gz;
    r.w += sp * (sb ? phi : phi - LE_XI * b.w);
    out[i] = r;
}

template <typename B>
hipError_t launch_le(B bodies, int64_t n, double box, double G, const double *Wt, double *partial,
                     int nblocks, double *coef, int sb, float4 *out32, double4 *out64,
                     hipStream_t st) {
    if (n <= 0)
        return hipSuccess;
    if (nblocks < 1 || nblocks > LE_MAX_BLOCKS || (out32 == nullptr) == (out64 == nullptr))
        return hipErrorInvalidValue;
    const int64_t chunk = (n + nblocks - 1) / nblocks;
    double *mu = partial + (size_t)nblocks * LE_NMON;
    hipLaunchKernelGGL(k_le_partials<B>, dim3(nblocks), dim3(256), 0, st, bodies, n, box, chunk,
                       partial);
    hipLaunchKernelGGL(k_le_reduce, dim3(LE_NMON), dim3(256), 0, st, partial, nblocks, mu);
    hipLaunchKernelGGL(k_le_coeffs, dim3(1), dim3(256), 0, st, mu, Wt, coef, sb);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (out32)
        hipLaunchKernelGGL((k_le_apply<B, float4>), grid, dim3(256), 0, st, bodies, n, box, G, coef,
                           sb, out32);
    else
        hipLaunchKernelGGL((k_le_apply<B, double4>), grid, dim3(256), 0, st, bodies, n, box, G,
                           coef, sb, out64);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ewald_le(const double4 *bpos, int64_t n, double box, double G, const double *Wt,
                           double *partial, int nblocks, double *coef, float4 *out32,
                           double4 *out64, hipStream_t st) {
    return launch_le(Bodies4{bpos}, n, box, G, Wt, partial, nblocks, coef, 0, out32, out64, st);
}

}  // namespace pgrav

// ---- the far field for caller device arrays: FixCubeEwaldLE as a GPU pass
namespace {

constexpr int LE_CACHE_DEVICES = 64;
std::mutex le_mu;
// per device: W transposed for nimage 0..MAX_NIMAGE, plain and subtract_background
double *le_wt[LE_CACHE_DEVICES][pgrav::MAX_NIMAGE + 1][2];

const double *device_wt(int nimage, int sb) {
    int dev = -1;
    if (!pgrav::check(hipGetDevice(&dev), "hipGetDevice"))
        return nullptr;
    if (dev < 0 || dev >= LE_CACHE_DEVICES) {
        pgrav::fail("pgrav_ewald_le_device: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(le_mu);
    double *&w = le_wt[dev][nimage][sb];
    if (!w) {
        std::vector<double> W, Wt((size_t)LE_NMON * LE_NMON);
        pgrav::le_matrix(nimage, W, sb != 0);
        for (int b = 0; b < LE_NMON; b++)
            for (int g = 0; g < LE_NMON; g++) Wt[(size_t)g * LE_NMON + b] = W[(size_t)b * LE_NMON + g];
        double *d = nullptr;
        if (!pgrav::check(hipMalloc(&d, sizeof(double) * Wt.size()), "hipMalloc") ||
            !pgrav::check(hipMemcpy(d, Wt.data(), sizeof(double) * Wt.size(), hipMemcpyHostToDevice),
                          "hipMemcpy"))
            return nullptr;
        w = d;  // kept until the process ends
    }
    return w;
}

}  // namespace

extern "C" int pgrav_ewald_le_device(const double *pos3, const double *mass, int64_t n, double box,
                                     int nimage, double G, int subtract_background, double *acc4,
                                     void *stream) {
    pgrav::clear_error();
    if (!pos3 || !mass || !acc4 || n < 0 || !(box > 0) || nimage < 0 ||
        nimage > pgrav::MAX_NIMAGE) {
        pgrav::fail("pgrav_ewald_le_device: bad arguments");
        return -1;
    }
    if (n == 0)
        return 0;
    try {
        const int sb = subtract_background ? 1 : 0;
        const double *Wt = device_wt(nimage, sb);
        if (!Wt)
            return -1;
        hipStream_t st = (hipStream_t)stream;
        const int nblocks = (int)std::min<int64_t>(2048, (n + 511) / 512);
        // stream-ordered scratch: the workgroup partials, the moments and the coefficients
        double *scr = nullptr;
        const size_t words = (size_t)LE_NMON * (nblocks + 1) + LE_NMON + 3;
        if (!pgrav::check(hipMallocAsync((void **)&scr, sizeof(double) * words, st), "hipMallocAsync"))
            return -1;
        double *partial = scr, *coef = scr + (size_t)LE_NMON * (nblocks + 1);
        const hipError_t e = pgrav::launch_le(pgrav::Bodies3{pos3, mass}, n, box, G, Wt, partial,
                                              nblocks, coef, sb, nullptr, (double4 *)acc4, st);
        const bool ok = pgrav::check(e, "far-field kernels");
        if (!pgrav::check(hipFreeAsync(scr, st), "hipFreeAsync") || !ok)
            return -1;
        return 0;
    } catch (const std::exception &ex) {
        pgrav::fail("pgrav_ewald_le_device: %s", ex.what());
        return -1;
    }
}
