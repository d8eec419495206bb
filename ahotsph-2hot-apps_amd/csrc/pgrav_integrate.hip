// Leapfrog kick-drift of the body table on the GPU (nln/integrate.c:16-59
// Integrate, :62-80 ConvertV, :82-135 CosmoIntegrate): one thread per body over
// the reference's strided float fields, the step in FP64 exactly as the C
// source orders and rounds it (no fused multiply-adds: every product and sum
// is rounded as the C expression is), kinetic and potential energy summed in
// FP64 per workgroup and then over the workgroups in a fixed order.
//
// HBM-bound: 44 B read and 24 B written per body (mass, pos, vel, acc, phi in;
// pos, vel out), plus the unused bytes of the caller's struct between them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/pgrav.h"
#include "pgrav_internal.h"

namespace pgrav {

namespace {

constexpr int64_t INTEGRATE_MAX_BLOCKS = 8192;  // workgroups of the grid-stride step

struct Table {
    const char *mass, *pos, *vel, *acc, *phi;
    char *opos, *ovel;
    int64_t in_stride, out_stride;
};

// kick factors: velocity at the position time (energy) and the step itself;
// COSMO: drift between comoving scalings a0 -> a1 (nln/integrate.c:118-124)
struct Step {
    double ksync, kick, drift, a0inv, a1, ke_scale;
};

__device__ __forceinline__ float ld(const char *p, int64_t i, int64_t stride, int k) {
    return ((const float *)(p + i * stride))[k];
}

// Grid-stride: a few thousand workgroups of 256 threads walk the table, each
// thread summing its bodies' energies; one LDS reduction per workgroup.
template <bool COSMO>
__global__ __launch_bounds__(256) void k_integrate(Table T, int64_t n, Step S, double2 *partial) {
#pragma clang fp contract(off)
    __shared__ double ske[256], spe[256];
    double ke = 0.0, pe = 0.0;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        const float mf = ld(T.mass, i, T.in_stride, 0), phif = ld(T.phi, i, T.in_stride, 0);
        const double m = mf;
        double vc2 = 0.0;
        float vout[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double v = ld(T.vel, i, T.in_stride, k), a = ld(T.acc, i, T.in_stride, k);
            const double vc = v + S.ksync * a;  // vcentered
            vc2 = vc2 + vc * vc;
            vout[k] = (float)(v + S.kick * a);
        }
        double kei = m * vc2;
        if (COSMO)
            kei = kei / S.ke_scale;  // / (a0 * a0)
        ke += kei;
        pe += (double)(mf * phif);  // *inmass * *inphi: a float product
        float pout[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double x = ld(T.pos, i, T.in_stride, k);
            if (COSMO) {
                float c = (float)(S.a0inv * x);                            // to comoving
                c = (float)((double)c + S.drift * (double)vout[k]);
                pout[k] = (float)(S.a1 * (double)c);                       // to physical
            } else {
                pout[k] = (float)(x + S.drift * (double)vout[k]);
            }
        }
        // stores after all loads: in and out may be the same table
        float *ov = (float *)(T.ovel + i * T.out_stride), *op = (float *)(T.opos + i * T.out_stride);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            ov[k] = vout[k];
            op[k] = pout[k];
        }
    }
    ske[threadIdx.x] = ke;
    spe[threadIdx.x] = pe;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ske[threadIdx.x] += ske[threadIdx.x + s];
            spe[threadIdx.x] += spe[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = make_double2(ske[0], spe[0]);
}

__global__ __launch_bounds__(256) void k_sum_partials(const double2 *partial, int64_t nb, double2 *out) {
    __shared__ double ske[256], spe[256];
    double ke = 0.0, pe = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) {
        ke += partial[b].x;
        pe += partial[b].y;
    }
    ske[threadIdx.x] = ke;
    spe[threadIdx.x] = pe;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ske[threadIdx.x] += ske[threadIdx.x + s];
            spe[threadIdx.x] += spe[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = make_double2(ske[0], spe[0]);
}

__global__ __launch_bounds__(256) void k_convert_v(char *vel, int64_t stride, int64_t n, double scale,
                                                   int to_physical) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float *v = (float *)(vel + i * stride);
    // nln/integrate.c:64: one_on_scale = (float)1.0 / scale, rounded to float
    const float inv = (float)(1.0 / scale);
#pragma unroll
    for (int k = 0; k < 3; k++)
        v[k] = to_physical ? v[k] * inv : (float)((double)v[k] * scale);
}

// the bytes of a strided table that hold the given fields (field k: len[k] floats)
void span(const void *const *f, const int *len, int nf, int64_t n, int64_t stride, const char **lo,
          const char **hi) {
    const char *a = (const char *)f[0], *b = (const char *)f[0] + 4 * len[0];
    for (int k = 1; k < nf; k++) {
        a = std::min(a, (const char *)f[k]);
        b = std::max(b, (const char *)f[k] + 4 * len[k]);
    }
    *lo = a;
    *hi = b + (n - 1) * stride;
}

// scratch: null, or nb + 1 double2 of device memory the caller provides (the
// workgroup partial sums, then the total); null allocates them for the call
int run(bool cosmo, const Table &T0, int64_t n, const Step &S, bool host, hipStream_t st,
        double *ke, double *pe, double2 *scratch = nullptr) {
    Table T = T0;
    char *din = nullptr, *dout = nullptr;
    double2 *partial = nullptr, *tot = nullptr;
    const int64_t nb = std::min<int64_t>((n + 255) / 256, INTEGRATE_MAX_BLOCKS);
    bool ok = true;
    if (scratch) {
        partial = scratch;
        tot = scratch + nb;
    } else {
        ok = check(hipMalloc(&partial, sizeof(double2) * nb), "hipMalloc") &&
             check(hipMalloc(&tot, sizeof(double2)), "hipMalloc");
    }
    const char *ilo = nullptr, *ihi = nullptr, *olo = nullptr, *ohi = nullptr;
    if (ok && host) {
        // copy the caller's tables (the spans their fields cover) to the device,
        // keeping offsets; one buffer when the output table overlaps the input
        const void *fi[5] = {T0.mass, T0.pos, T0.vel, T0.acc, T0.phi}, *fo[2] = {T0.opos, T0.ovel};
        const int li[5] = {1, 3, 3, 3, 1}, lo3[2] = {3, 3};
        span(fi, li, 5, n, T0.in_stride, &ilo, &ihi);
        span(fo, lo3, 2, n, T0.out_stride, &olo, &ohi);
        const bool overlap = olo < ihi && ilo < ohi;
        if (overlap) {
            ilo = std::min(ilo, olo);
            ihi = std::max(ihi, ohi);
            olo = ilo;
            ohi = ihi;
        }
        ok = check(hipMalloc(&din, ihi - ilo), "hipMalloc") &&
             check(hipMemcpyAsync(din, ilo, ihi - ilo, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        if (ok && !overlap)
            ok = check(hipMalloc(&dout, ohi - olo), "hipMalloc") &&
                 check(hipMemcpyAsync(dout, olo, ohi - olo, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        if (overlap)
            dout = din;
        auto in = [&](const char *p) { return din + (p - ilo); };
        auto out = [&](char *p) { return dout + (p - olo); };
        T.mass = in(T0.mass);
        T.pos = in(T0.pos);
        T.vel = in(T0.vel);
        T.acc = in(T0.acc);
        T.phi = in(T0.phi);
        T.opos = out(T0.opos);
        T.ovel = out(T0.ovel);
    }
    if (ok) {
        if (cosmo)
            hipLaunchKernelGGL(k_integrate<true>, dim3((unsigned)nb), dim3(256), 0, st, T, n, S, partial);
        else
            hipLaunchKernelGGL(k_integrate<false>, dim3((unsigned)nb), dim3(256), 0, st, T, n, S, partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, nb, tot);
        ok = check(hipGetLastError(), "integrator kernels");
    }
    double2 h = make_double2(0, 0);
    if (ok)
        ok = check(hipMemcpyAsync(&h, tot, sizeof(h), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (ok && host)
        ok = check(hipMemcpyAsync((void *)olo, dout, ohi - olo, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (ok)
        ok = check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (void *p : {(void *)din, (void *)(dout != din ? dout : nullptr)})
        if (p)
            (void)hipFree(p);
    if (!scratch)
        for (void *p : {(void *)partial, (void *)tot})
            if (p)
                (void)hipFree(p);
    if (!ok)
        return -1;
    // nln/integrate.c:56-57: kep = 0.5 * ke, pep = 0.5 * pe
    *ke = 0.5 * h.x;
    *pe = 0.5 * h.y;
    return 0;
}

int integrate_common(bool cosmo, const float *inmass, const float *inpos, const float *invel,
                     const float *inacc, const float *inphi, int instride, float *outpos,
                     float *outvel, int outstride, int n, const Step &S, bool host, void *stream,
                     double *kep, double *pep, double2 *scratch = nullptr) {
    clear_error();
    if (!inmass || !inpos || !invel || !inacc || !inphi || !outpos || !outvel || !kep || !pep ||
        n < 0 || instride <= 0 || outstride <= 0 || instride % 4 || outstride % 4) {
        fail("integrate: bad arguments (strides must be positive multiples of sizeof(float), "
             "nln/integrate.c:35-36)");
        return -1;
    }
    if (n == 0) {
        *kep = 0.0;
        *pep = 0.0;
        return 0;
    }
    Table T{(const char *)inmass, (const char *)inpos, (const char *)invel, (const char *)inacc,
            (const char *)inphi, (char *)outpos, (char *)outvel, instride, outstride};
    return run(cosmo, T, n, S, host, (hipStream_t)stream, kep, pep, scratch);
}

// ---- the entry point that allocates nothing: its workgroup partial sums and
// total live in this file-scope array, one copy per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel (on the host it is
// only a stand-in for the symbol); it passes the device address that
// hipGetSymbolAddress returns, looked up once per device.
constexpr int64_t NOALLOC_MAX_BODIES = int64_t(1) << 27;  // 134,217,728 bodies
constexpr int64_t NOALLOC_SLOTS = NOALLOC_MAX_BODIES / 256 + 1;  // 8 MiB + 16 B
static __device__ __align__(256) double2 scratch[NOALLOC_SLOTS];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double2 *scratch_ptr[MAX_DEVICES];

double2 *device_scratch() {
    int dev = -1;
    if (!check(hipGetDevice(&dev), "hipGetDevice"))
        return nullptr;
    if (dev < 0 || dev >= MAX_DEVICES) {
        fail("pgrav_integrate_device_noalloc: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        if (!check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress"))
            return nullptr;
        scratch_ptr[dev] = (double2 *)p;
    }
    return scratch_ptr[dev];
}

}  // namespace

}  // namespace pgrav

using namespace pgrav;

extern "C" {

int pgrav_integrate(const float *inmass, const float *inpos, const float *invel, const float *inacc,
                    const float *inphi, int instride, float *outpos, float *outvel, int outstride,
                    int n, double dt, double dtv, double *tpos, double *tvel, double *kep,
                    double *pep) {
    if (!tpos || !tvel) {
        fail("pgrav_integrate: null time");
        return -1;
    }
    Step S{*tpos - *tvel, dtv, dt, 1.0, 1.0, 1.0};
    if (integrate_common(false, inmass, inpos, invel, inacc, inphi, instride, outpos, outvel,
                         outstride, n, S, true, nullptr, kep, pep))
        return -1;
    *tvel += dtv;
    *tpos += dt;
    return 0;
}

int pgrav_integrate_device(const float *inmass, const float *inpos, const float *invel,
                           const float *inacc, const float *inphi, int instride, float *outpos,
                           float *outvel, int outstride, int n, double dt, double dtv, double *tpos,
                           double *tvel, double *kep, double *pep, void *stream) {
    if (!tpos || !tvel) {
        fail("pgrav_integrate_device: null time");
        return -1;
    }
    Step S{*tpos - *tvel, dtv, dt, 1.0, 1.0, 1.0};
    if (integrate_common(false, inmass, inpos, invel, inacc, inphi, instride, outpos, outvel,
                         outstride, n, S, false, stream, kep, pep))
        return -1;
    *tvel += dtv;
    *tpos += dt;
    return 0;
}

int pgrav_integrate_device_noalloc(const float *inmass, const float *inpos, const float *invel,
                                   const float *inacc, const float *inphi, int instride,
                                   float *outpos, float *outvel, int outstride, int n, double dt,
                                   double dtv, double *tpos, double *tvel, double *kep, double *pep,
                                   void *stream) {
    clear_error();
    if (!tpos || !tvel) {
        fail("pgrav_integrate_device_noalloc: null time");
        return -1;
    }
    if ((int64_t)n > NOALLOC_MAX_BODIES) {
        fail("pgrav_integrate_device_noalloc: n exceeds the 134217728 bodies its static scratch "
             "holds (use pgrav_integrate_device)");
        return -1;
    }
    double2 *ws = device_scratch();
    if (!ws)
        return -1;
    Step S{*tpos - *tvel, dtv, dt, 1.0, 1.0, 1.0};
    if (integrate_common(false, inmass, inpos, invel, inacc, inphi, instride, outpos, outvel,
                         outstride, n, S, false, stream, kep, pep, ws))
        return -1;
    *tvel += dtv;
    *tpos += dt;
    return 0;
}

int pgrav_cosmo_integrate_device(const float *inmass, const float *inpos, const float *invel,
                                 const float *inacc, const float *inphi, int instride,
                                 float *outpos, float *outvel, int outstride, int n, double a0,
                                 double a1, double dt_kick_topos, double dt_kick, double dt_drift,
                                 double *kep, double *pep, void *stream) {
    if (!(a0 > 0) || !(a1 > 0)) {
        fail("pgrav_cosmo_integrate_device: scale factors must be positive");
        return -1;
    }
    Step S{dt_kick_topos, dt_kick, dt_drift, 1.0 / a0, a1, a0 * a0};
    return integrate_common(true, inmass, inpos, invel, inacc, inphi, instride, outpos, outvel,
                            outstride, n, S, false, stream, kep, pep);
}

int pgrav_convert_v_device(float *vel, int stride, int n, double scale, int to_physical,
                           void *stream) {
    clear_error();
    if (!vel || n < 0 || stride <= 0 || stride % 4 || !(scale != 0)) {
        fail("pgrav_convert_v_device: bad arguments");
        return -1;
    }
    if (n == 0)
        return 0;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_convert_v, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (char *)vel,
                       (int64_t)stride, (int64_t)n, scale, to_physical);
    return check(hipGetLastError(), "ConvertV kernel") && check(hipStreamSynchronize(st), "hipStreamSynchronize")
               ? 0 : -1;
}

}  // extern "C"
