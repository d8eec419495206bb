// Kernels behind the drop-in entry points of nln/gravcuda.h (grav_mn_CUDA,
// grav_mns_CUDA, grav_mnss_CUDA, grav_qns_CUDA, grav_qnss_CUDA).
//
// MI355X mapping: one wave64 = 64 sinks, one lane per sink.  Sources that every
// lane of a wave shares (the M x N kernels) are read through wave-uniform
// addresses, so the compiler emits scalar loads into SGPRs and the whole wave
// consumes each source record as a broadcast operand: no LDS traffic and no
// per-lane gathers.  Long source lists are split over blockIdx.y so that one
// call fills the chip; partial results meet in the device accumulator with a
// float atomic add, exactly as the reference's concurrent streams do
// (nln/pgrav.cu:207-210).
#include "pgrav_device.h"
#include "pgrav_internal.h"

using namespace pgrav;

// nln/pgrav.cu:16-47, 66: SoA-8 source field k of source i.
__device__ __forceinline__ float soa8(const float *__restrict__ f, int i, int k, int sz) {
    return f[(i >> 3) * 8 * sz + k * 8 + (i & 7)];
}

// SoA-8 reference records -> (pos4, prepared record).  Zero-mass padding
// entries (nln/mac.c:486-520 pads every list to a multiple of NSSE with zero
// records) are parked at 1e8 so they contribute exactly zero even for a sink
// at the origin (where the reference would form 0 * inf).  1e8 keeps x^4 finite
// in FP32, so no inf - inf can appear in the order-4 basis.
// the order is a template parameter so that the record stays in registers
template <int ORDER>
__device__ __forceinline__ void prep_source(const float *__restrict__ f, int i, int sz,
                                            int dipole, float4 *__restrict__ pos4,
                                            float *__restrict__ mom) {
    constexpr int NF = ORDER == 0 ? 4 : (ORDER == 2 ? 13 : 29);
    constexpr int MSTRIDE = ORDER == 4 ? PREC4 : PREC2;
    float q[29];
#pragma unroll
    for (int k = 0; k < 29; k++) q[k] = k < NF ? soa8(f, i, k, sz) : 0.0f;
    float4 p = make_float4(q[1], q[2], q[3], q[0]);
    bool zero = q[0] == 0.0f;
#pragma unroll
    for (int k = 4; k < NF; k++) zero = zero && q[k] == 0.0f;
    if (zero)
        p = make_float4(1e8f, 1e8f, 1e8f, 0.0f);
    pos4[i] = p;
    if (ORDER > 0) {
        float rec[PREC4];
        prep_record(q, ORDER, dipole, rec);
#pragma unroll
        for (int k = 0; k < MSTRIDE; k++) mom[(size_t)i * MSTRIDE + k] = rec[k];
    }
}

template <int ORDER>
__global__ __launch_bounds__(128) void k_prep_soa8(const float *__restrict__ f, int source_n,
                                                   int sz, int dipole, float4 *__restrict__ pos4,
                                                   float *__restrict__ mom) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < source_n)
        prep_source<ORDER>(f, i, sz, dipole, pos4, mom);
}

// AoS hexacell table (nln/physics.h:140-157, byte stride `stride`) -> prepared
// order-2 table for pQQ1.  Field offsets in floats: mass 0, pos 1..3, bmax 5,
// qx..qz 12..14, qxx..qyz 15..19.
__global__ void k_prep_hexacell(const float *__restrict__ tab, int stride_floats, int64_t n,
                                float4 *__restrict__ pos4, float *__restrict__ mom) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float *c = tab + i * stride_floats;
    float q[29] = {0};
    q[0] = c[0]; q[1] = c[1]; q[2] = c[2]; q[3] = c[3]; q[4] = c[5];
    for (int k = 0; k < 8; k++) q[5 + k] = c[12 + k];
    pos4[i] = make_float4(q[1], q[2], q[3], q[0]);
    float rec[PREC4];
    prep_record(q, 2, 1, rec);
    for (int k = 0; k < PREC2; k++) mom[i * PREC2 + k] = rec[k];
}

__device__ __forceinline__ void flush_acc(float *__restrict__ acc, const Acc4 &a) {
    atomicAdd(acc + 0, a.x);
    atomicAdd(acc + 1, a.y);
    atomicAdd(acc + 2, a.z);
    atomicAdd(acc + 3, a.w);
}

// Sources [i0, i1) of a prepared list against one sink.  Every lane reads the
// same record, so the loads are scalar (wave-uniform addresses).
template <int ORDER, int QSIGN, int SOFT>
__device__ __forceinline__ void mn_range(float px, float py, float pz,
                                         const float4 *__restrict__ pos4,
                                         const float *__restrict__ mom, int i0, int i1, float e,
                                         Acc4 &a) {
    const float eps2k1 = SOFT == SOFT_K1 ? 1.0f / (e * e) : 0.0f;
    for (int i = i0; i < i1; i++) {
        const float4 q = pos4[i];
        const float x = px - q.x, y = py - q.y, z = pz - q.z;
        if (ORDER == 0) {
            if (SOFT == SOFT_K1)
                k1_interact(q.w, x, y, z, e, eps2k1, a);
            else
                mono_interact(q.w, x, y, z, 0.0f, a);
        } else {
            const float *M = mom + (size_t)i * (ORDER == 4 ? PREC4 : PREC2);
            cell_interact<ORDER, true, QSIGN>(M, q.w, x, y, z, 0.0f, a);
        }
    }
}

// pH / pQ / pM / pM_sK1 (nln/pgrav.cu:49-468): m sinks x n sources.
// ORDER 0 with SOFT == SOFT_K1 is pM_sK1 (e = 1/eps).
template <int ORDER, int QSIGN, int SOFT>
__global__ __launch_bounds__(64) void k_mn(const float *__restrict__ sinkpos,
                                           float *__restrict__ acc, int m,
                                           const float4 *__restrict__ pos4,
                                           const float *__restrict__ mom, int n, int chunk,
                                           float e) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    const int i0 = blockIdx.y * chunk;
    const int i1 = min(n, i0 + chunk);
    const bool live = s < m;
    const int sc = live ? s : 0;  // dead lanes compute on a valid sink, never store
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    mn_range<ORDER, QSIGN, SOFT>(sinkpos[3 * sc], sinkpos[3 * sc + 1], sinkpos[3 * sc + 2], pos4,
                                 mom, i0, i1, e, a);
    if (live)
        flush_acc(acc + 4 * s, a);
}

// Bodies idx[j0..j1) of a pMM_sK1 list against one sink, staged through LDS 64
// at a time (one coalesced load per tile, however short the segments are).
// Every lane of the workgroup (one wave) takes part in the staging.
__device__ __forceinline__ void mns_range(float px, float py, float pz,
                                          const uint32_t *__restrict__ idx, int j0, int j1,
                                          const float *__restrict__ srcpos, float mmass,
                                          float eps_inv, float3 *tile, Acc4 &a) {
    const float eps2 = 1.0f / (eps_inv * eps_inv);
    for (int t0 = j0; t0 < j1; t0 += 64) {
        const int nt = min(64, j1 - t0);
        if ((int)threadIdx.x < nt) {
            const float *q = srcpos + 3 * (size_t)idx[t0 + threadIdx.x];
            tile[threadIdx.x] = make_float3(q[0], q[1], q[2]);
        }
        __syncthreads();
#pragma unroll 4
        for (int j = 0; j < nt; j++) {
            const float3 q = tile[j];
            k1_interact(mmass, px - q.x, py - q.y, pz - q.z, eps_inv, eps2, a);
        }
        __syncthreads();
    }
}

// pMM_sK1 (nln/pgrav.cu:575-637): every sink against the bodies of a
// segment list into the sink table.  The host expands the segments into body
// indices; grid.y splits that list so that a call runs on several waves.
__global__ __launch_bounds__(64) void k_mns(const float *__restrict__ sinkpos,
                                            float *__restrict__ acc, int m,
                                            const uint32_t *__restrict__ idx, int nidx, int chunk,
                                            const float *__restrict__ srcpos, float mmass,
                                            float eps_inv) {
    __shared__ float3 tile[64];
    const int s = blockIdx.x * 64 + threadIdx.x;
    const bool live = s < m;
    const int sc = live ? s : 0;
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    const int j0 = blockIdx.y * chunk, j1 = min(nidx, j0 + chunk);
    mns_range(sinkpos[3 * sc], sinkpos[3 * sc + 1], sinkpos[3 * sc + 2], idx, j0, j1, srcpos,
              mmass, eps_inv, tile, a);
    if (live && j0 < j1)
        flush_acc(acc + 4 * s, a);
}

// The batched queue kernels below give each sink one lane and cut its source list
// into `split` contiguous slices, W per workgroup (one wave each): a flush of a
// few hundred sinks then still launches thousands of waves instead of a handful
// of latency-bound ones.  The W slices of a workgroup are summed in LDS in wave
// order and added once into the accumulator.
template <int W>
__device__ __forceinline__ void sum_slices(float4 (*part)[64], int lane, int w, bool live,
                                           const Acc4 &a, float *acc) {
    part[w][lane] = make_float4(a.x, a.y, a.z, a.w);
    __syncthreads();
    if (w == 0 && live) {
        Acc4 t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < W; k++) {
            const float4 v = part[k][lane];
            t.x += v.x;
            t.y += v.y;
            t.z += v.z;
            t.w += v.w;
        }
        flush_acc(acc, t);
    }
}

// pMM1_mnss_sK1 (nln/pgrav.cu:713-790): sink i uses segment run ss_seg[ss_index[i]]
// of the segment array; slice k of a sink covers bodies [k c, (k + 1) c) of the
// run's bodies in segment order, c = ceil(bodies / split).  The count check of
// the reference (its device printf, nln/pgrav.cu:779-789) is made once per sink.
template <int W>
__global__ __launch_bounds__(64 * W) void k_mnss(const float *__restrict__ sinkpos,
                                                 float *__restrict__ acc, int m,
                                                 const uint16_t *__restrict__ ss_index,
                                                 const uint32_t *__restrict__ ss_seg, int ss_len,
                                                 const uint32_t *__restrict__ s_seg, int s_seg_len,
                                                 const float *__restrict__ srcpos,
                                                 const int *__restrict__ s_source_n, float mmass,
                                                 float eps_inv, int *__restrict__ errflag,
                                                 int split) {
    __shared__ float4 part[W][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + lane;
    const int slice = blockIdx.y * W + w;
    bool live = s < m;
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    if (live) {
        const int si = ss_index[s];
        if (si >= ss_len || ss_seg[2 * si] + ss_seg[2 * si + 1] > (uint32_t)s_seg_len) {
            if (slice == 0)
                atomicAdd(errflag, 1);
            live = false;
        } else {
            const uint32_t *seg = s_seg + 2 * (size_t)ss_seg[2 * si];
            const int seg_len = (int)ss_seg[2 * si + 1];
            uint32_t total = 0;
            for (int g = 0; g < seg_len; g++) total += seg[2 * g + 1];
            if (slice == 0 && (int)total != s_source_n[si])
                atomicAdd(errflag, 1);
            const uint32_t chunk = (total + split - 1) / split;
            const uint32_t j0 = (uint32_t)min((uint64_t)total, (uint64_t)slice * chunk);
            const uint32_t j1 = min(total, j0 + chunk);
            const float px = sinkpos[3 * s], py = sinkpos[3 * s + 1], pz = sinkpos[3 * s + 2];
            const float eps2 = 1.0f / (eps_inv * eps_inv);
            uint32_t off = 0;
            for (int g = 0; g < seg_len && off < j1; g++) {
                const uint32_t base = seg[2 * g], len = seg[2 * g + 1];
                const uint32_t lo = max(j0, off), hi = min(j1, off + len);
                for (uint32_t j = lo; j < hi; j++) {
                    const float *q = srcpos + 3 * (size_t)(base + (j - off));
                    k1_interact(mmass, px - q[0], py - q[1], pz - q[2], eps_inv, eps2, a);
                }
                off += len;
            }
        }
    }
    sum_slices<W>(part, lane, w, live, a, acc + 4 * (size_t)(live ? s : 0));
}

// pQ1_mnss (nln/pgrav.cu:293-368): sink i uses run seg[i] of the packed buffer;
// slice k of a sink covers [k c, (k + 1) c) of its run, c = ceil(length / split).
template <int QSIGN, int W>
__global__ __launch_bounds__(64 * W) void k_qnss(const float *__restrict__ sinkpos,
                                                 float *__restrict__ acc, int m,
                                                 const uint32_t *__restrict__ seg,
                                                 const float4 *__restrict__ pos4,
                                                 const float *__restrict__ mom, int split) {
    __shared__ float4 part[W][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s = blockIdx.x * 64 + lane;
    const bool live = s < m;
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    if (live) {
        const uint32_t base = seg[2 * s], len = seg[2 * s + 1];
        const uint32_t chunk = (len + split - 1) / split;
        const uint32_t j0 = (uint32_t)min((uint64_t)len, (uint64_t)(blockIdx.y * W + w) * chunk);
        const uint32_t j1 = min(len, j0 + chunk);
        const float px = sinkpos[3 * s], py = sinkpos[3 * s + 1], pz = sinkpos[3 * s + 2];
        for (uint32_t j = j0; j < j1; j++) {
            const float4 q = pos4[base + j];
            cell_interact<2, true, QSIGN>(mom + (size_t)(base + j) * PREC2, q.w, px - q.x,
                                          py - q.y, pz - q.z, 0.0f, a);
        }
    }
    sum_slices<W>(part, lane, w, live, a, acc + 4 * (size_t)(live ? s : 0));
}

// Cells idx[i0..i1) of the WalkInitSrcCUDA table against one sink (scalar loads:
// every lane reads the same index and record).
template <int QSIGN>
__device__ __forceinline__ void qns_range(float px, float py, float pz,
                                          const int *__restrict__ idx, int i0, int i1,
                                          const float4 *__restrict__ pos4,
                                          const float *__restrict__ mom, Acc4 &a) {
    for (int i = i0; i < i1; i++) {
        const int c = idx[i];
        const float4 q = pos4[c];
        cell_interact<2, true, QSIGN>(mom + (size_t)c * PREC2, q.w, px - q.x, py - q.y, pz - q.z,
                                      0.0f, a);
    }
}

// pQQ1 (nln/pgrav.cu:498-572) with the defects of A5 fixed: bound on the
// sink count and idx[i] selecting the record.
template <int QSIGN>
__global__ __launch_bounds__(64) void k_qns(const float *__restrict__ sinkpos,
                                            float *__restrict__ acc, int m,
                                            const int *__restrict__ idx, int source_n,
                                            const float4 *__restrict__ pos4,
                                            const float *__restrict__ mom) {
    const int s = blockIdx.x * 64 + threadIdx.x;
    const bool live = s < m;
    const int sc = live ? s : 0;
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    qns_range<QSIGN>(sinkpos[3 * sc], sinkpos[3 * sc + 1], sinkpos[3 * sc + 2], idx, 0, source_n,
                     pos4, mom, a);
    if (live)
        flush_acc(acc + 4 * s, a);
}

// ------------------------------------------------------- deferred-call batches
// One upload of many calls' lists and one launch per kind, instead of a copy and
// one or two launches per call.  Each workgroup (one wave, 64 sinks) takes one
// task: a call, a block of its sinks and a range of its sources; the call and
// task records are wave-uniform, so they come through scalar loads.
template <int ORDER>
__global__ __launch_bounds__(128) void k_prep_batch(const float *__restrict__ data,
                                                    const DeferredCall *__restrict__ calls,
                                                    const int *__restrict__ block_call,
                                                    float4 *__restrict__ pos4,
                                                    float *__restrict__ mom) {
    constexpr int MSTRIDE = ORDER == 4 ? PREC4 : PREC2;
    const DeferredCall c = calls[block_call[blockIdx.x]];
    const int i = (int)(blockIdx.x - c.prep / 128) * 128 + threadIdx.x;
    if (i < c.n)
        prep_source<ORDER>(data + c.src, i, c.sz, 1, pos4 + c.prep,
                           mom + (size_t)c.prep * MSTRIDE);
}

template <int ORDER, int QSIGN, int SOFT>
__global__ __launch_bounds__(64) void k_mn_batch(const float *__restrict__ sinkpos,
                                                 float *__restrict__ acc,
                                                 const DeferredCall *__restrict__ calls,
                                                 const DeferredTask *__restrict__ tasks,
                                                 const float4 *__restrict__ pos4,
                                                 const float *__restrict__ mom) {
    constexpr int MSTRIDE = ORDER == 4 ? PREC4 : PREC2;
    const DeferredTask t = tasks[blockIdx.x];
    const DeferredCall c = calls[t.x];
    const int s = t.y + threadIdx.x;
    const bool live = s < c.m;
    const float *p = sinkpos + 3 * (c.sink + (live ? s : 0));
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    mn_range<ORDER, QSIGN, SOFT>(p[0], p[1], p[2], pos4 + c.prep, mom + (size_t)c.prep * MSTRIDE,
                                 t.z, t.w, c.e, a);
    if (live)
        flush_acc(acc + 4 * (c.sink + s), a);
}

__global__ __launch_bounds__(64) void k_mns_batch(const float *__restrict__ sinkpos,
                                                  float *__restrict__ acc,
                                                  const DeferredCall *__restrict__ calls,
                                                  const DeferredTask *__restrict__ tasks,
                                                  const uint32_t *__restrict__ data) {
    __shared__ float3 tile[64];
    const DeferredTask t = tasks[blockIdx.x];
    const DeferredCall c = calls[t.x];
    const int s = t.y + threadIdx.x;
    const bool live = s < c.m;
    const float *p = sinkpos + 3 * (c.sink + (live ? s : 0));
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    mns_range(p[0], p[1], p[2], data + c.src, t.z, t.w, sinkpos, c.mass, c.e, tile, a);
    if (live)
        flush_acc(acc + 4 * (c.sink + s), a);
}

template <int QSIGN>
__global__ __launch_bounds__(64) void k_qns_batch(const float *__restrict__ sinkpos,
                                                  float *__restrict__ acc,
                                                  const DeferredCall *__restrict__ calls,
                                                  const DeferredTask *__restrict__ tasks,
                                                  const int *__restrict__ data,
                                                  const float4 *__restrict__ pos4,
                                                  const float *__restrict__ mom) {
    const DeferredTask t = tasks[blockIdx.x];
    const DeferredCall c = calls[t.x];
    const int s = t.y + threadIdx.x;
    const bool live = s < c.m;
    const float *p = sinkpos + 3 * (c.sink + (live ? s : 0));
    Acc4 a = {0.f, 0.f, 0.f, 0.f};
    qns_range<QSIGN>(p[0], p[1], p[2], data + c.src, t.z, t.w, pos4, mom, a);
    if (live)
        flush_acc(acc + 4 * (c.sink + s), a);
}

// ---------------------------------------------------------------- launchers
namespace pgrav {

hipError_t launch_prep_soa8(const float *f, int source_n, int sz, int order, float4 *pos4,
                            float *mom, hipStream_t st) {
    if (source_n <= 0)
        return hipSuccess;
    const dim3 grid((source_n + 127) / 128);
    if (order == 4)
        hipLaunchKernelGGL(k_prep_soa8<4>, grid, dim3(128), 0, st, f, source_n, sz, 1, pos4, mom);
    else if (order == 2)
        hipLaunchKernelGGL(k_prep_soa8<2>, grid, dim3(128), 0, st, f, source_n, sz, 1, pos4, mom);
    else
        hipLaunchKernelGGL(k_prep_soa8<0>, grid, dim3(128), 0, st, f, source_n, sz, 1, pos4, mom);
    return hipGetLastError();
}

hipError_t launch_prep_hexacell(const float *tab, int stride_floats, int64_t n, float4 *pos4,
                                float *mom, hipStream_t st) {
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_prep_hexacell, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, tab,
                       stride_floats, n, pos4, mom);
    return hipGetLastError();
}

hipError_t launch_mn(int order, int soft, int qsign, const float *sinkpos, float *acc, int m,
                     const float4 *pos4, const float *mom, int n, float e, hipStream_t st) {
    if (m <= 0 || n <= 0)
        return hipSuccess;
    const int sink_blocks = (m + 63) / 64;
    // Split the source list so that one call launches >= ~1024 waves when the
    // list is long; each wave keeps >= 32 sources to amortise the atomics (a
    // call of one sink group otherwise runs on a handful of waves for tens of
    // microseconds while the rest of the GPU idles).
    int split = 1;
    while (sink_blocks * split < 1024 && n / (split * 2) >= 32) split *= 2;
    const int chunk = (n + split - 1) / split;
    dim3 grid(sink_blocks, split), block(64);
#define MN(O, Q, S) hipLaunchKernelGGL((k_mn<O, Q, S>), grid, block, 0, st, sinkpos, acc, m, pos4, mom, n, chunk, e)
    if (order == 0) {
        if (soft == SOFT_K1) MN(0, 1, SOFT_K1);
        else MN(0, 1, SOFT_NONE);
    } else if (order == 2) {
        if (qsign < 0) MN(2, -1, SOFT_NONE);
        else MN(2, 1, SOFT_NONE);
    } else {
        if (qsign < 0) MN(4, -1, SOFT_NONE);
        else MN(4, 1, SOFT_NONE);
    }
#undef MN
    return hipGetLastError();
}

hipError_t launch_mns(const float *sinkpos, float *acc, int m, const uint32_t *idx, int nidx,
                      const float *srcpos, float mmass, float eps_inv, hipStream_t st) {
    if (m <= 0 || nidx <= 0)
        return hipSuccess;
    // >= 256 bodies per wave, up to 64 waves per call
    const int split = std::max(1, std::min(64, nidx / 256));
    const int chunk = (nidx + split - 1) / split;
    hipLaunchKernelGGL(k_mns, dim3((m + 63) / 64, split), dim3(64), 0, st, sinkpos, acc, m, idx,
                       nidx, chunk, srcpos, mmass, eps_inv);
    return hipGetLastError();
}

// Slices per sink for the queue kernels: at least ~2048 waves per call while
// a slice keeps >= `minslice` sources of the longest list; W slices per workgroup.
static int queue_split(int m, int64_t maxlen, int minslice) {
    const int64_t blocks = (m + 63) / 64;
    int split = 1;
    while (split < 256 && blocks * split < 2048 && maxlen / (2 * split) >= minslice) split *= 2;
    return split;
}

hipError_t launch_mnss(const float *sinkpos, float *acc, int m, const uint16_t *ss_index,
                       const uint32_t *ss_seg, int ss_len, const uint32_t *seg, int seg_len,
                       const float *srcpos, const int *source_n, int64_t max_source_n,
                       float mmass, float eps_inv, int *errflag, hipStream_t st) {
    if (m <= 0)
        return hipSuccess;
    const int split = queue_split(m, max_source_n, 32);
    const unsigned bx = (unsigned)((m + 63) / 64);
#define MNSS(W)                                                                                   \
    hipLaunchKernelGGL((k_mnss<W>), dim3(bx, split / (W)), dim3(64 * (W)), 0, st, sinkpos, acc, m, \
                       ss_index, ss_seg, ss_len, seg, seg_len, srcpos, source_n, mmass, eps_inv,  \
                       errflag, split)
    if (split >= 4)
        MNSS(4);
    else if (split == 2)
        MNSS(2);
    else
        MNSS(1);
#undef MNSS
    return hipGetLastError();
}

hipError_t launch_qnss(int qsign, const float *sinkpos, float *acc, int m, const uint32_t *seg,
                       int64_t max_len, const float4 *pos4, const float *mom, hipStream_t st) {
    if (m <= 0)
        return hipSuccess;
    const int split = queue_split(m, max_len, 16);
    const unsigned bx = (unsigned)((m + 63) / 64);
#define QNSS(Q, W)                                                                                \
    hipLaunchKernelGGL((k_qnss<Q, W>), dim3(bx, split / (W)), dim3(64 * (W)), 0, st, sinkpos, acc, \
                       m, seg, pos4, mom, split)
    if (qsign < 0) {
        if (split >= 4) QNSS(-1, 4);
        else if (split == 2) QNSS(-1, 2);
        else QNSS(-1, 1);
    } else {
        if (split >= 4) QNSS(1, 4);
        else if (split == 2) QNSS(1, 2);
        else QNSS(1, 1);
    }
#undef QNSS
    return hipGetLastError();
}

hipError_t launch_qns(int qsign, const float *sinkpos, float *acc, int m, const int *idx,
                      int source_n, const float4 *pos4, const float *mom, hipStream_t st) {
    if (m <= 0)
        return hipSuccess;
    if (qsign < 0)
        hipLaunchKernelGGL(k_qns<-1>, dim3((m + 63) / 64), dim3(64), 0, st, sinkpos, acc, m, idx,
                           source_n, pos4, mom);
    else
        hipLaunchKernelGGL(k_qns<1>, dim3((m + 63) / 64), dim3(64), 0, st, sinkpos, acc, m, idx,
                           source_n, pos4, mom);
    return hipGetLastError();
}

hipError_t launch_mn_batch(int order, int soft, int qsign, const float *data,
                           const DeferredCall *calls, const int *block_call, int nblocks,
                           const DeferredTask *tasks, int ntasks, float4 *pos4, float *mom,
                           const float *sinkpos, float *acc, hipStream_t st) {
    if (nblocks <= 0 || ntasks <= 0)
        return hipSuccess;
#define PREP(O)                                                                                 \
    hipLaunchKernelGGL(k_prep_batch<O>, dim3(nblocks), dim3(128), 0, st, data, calls, block_call, \
                       pos4, mom)
#define MNB(O, Q, S)                                                                            \
    hipLaunchKernelGGL((k_mn_batch<O, Q, S>), dim3(ntasks), dim3(64), 0, st, sinkpos, acc, calls, \
                       tasks, pos4, mom)
    if (order == 0) {
        PREP(0);
        if (soft == SOFT_K1) MNB(0, 1, SOFT_K1);
        else MNB(0, 1, SOFT_NONE);
    } else if (order == 2) {
        PREP(2);
        if (qsign < 0) MNB(2, -1, SOFT_NONE);
        else MNB(2, 1, SOFT_NONE);
    } else {
        PREP(4);
        if (qsign < 0) MNB(4, -1, SOFT_NONE);
        else MNB(4, 1, SOFT_NONE);
    }
#undef PREP
#undef MNB
    return hipGetLastError();
}

hipError_t launch_mns_batch(const uint32_t *data, const DeferredCall *calls,
                            const DeferredTask *tasks, int ntasks, const float *sinkpos,
                            float *acc, hipStream_t st) {
    if (ntasks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_mns_batch, dim3(ntasks), dim3(64), 0, st, sinkpos, acc, calls, tasks,
                       data);
    return hipGetLastError();
}

hipError_t launch_qns_batch(int qsign, const int *data, const DeferredCall *calls,
                            const DeferredTask *tasks, int ntasks, const float4 *pos4,
                            const float *mom, const float *sinkpos, float *acc, hipStream_t st) {
    if (ntasks <= 0)
        return hipSuccess;
    if (qsign < 0)
        hipLaunchKernelGGL(k_qns_batch<-1>, dim3(ntasks), dim3(64), 0, st, sinkpos, acc, calls,
                           tasks, data, pos4, mom);
    else
        hipLaunchKernelGGL(k_qns_batch<1>, dim3(ntasks), dim3(64), 0, st, sinkpos, acc, calls,
                           tasks, data, pos4, mom);
    return hipGetLastError();
}

}  // namespace pgrav
