// Internal (non-exported) declarations shared by the libpgrav translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgrav {

// One deferred call of the drop-in path (grav_mn_CUDA, pinteractCUDA,
// grav_mns_CUDA, grav_qns_CUDA) as the batched kernels read it.
struct DeferredCall {
    int64_t sink;  // first sink (body index)
    int64_t src;   // 4-byte word offset of the call's list in the batch data
    int m, n;      // sinks; sources (M x N, pQQ1) or bodies (pMM_sK1)
    int sz;        // M x N: floats per SoA-8 source record
    int prep;      // M x N: first prepared record of the call (a multiple of 128)
    float e;       // M x N: the call's e (pM_sK1: 1/eps); pMM_sK1: 1/eps
    float mass;    // pMM_sK1: the body mass
};
// (call, first sink, first source, end source): one workgroup of a batched kernel
typedef int4 DeferredTask;

// Error reporting (pgrav_abi.hip).  The reference aborts the job on every
// runtime error (Error() in libsw); pgrav_set_error_mode(1) turns that into a
// recorded message that pgrav_last_error() returns, for callers that want to
// handle it.
void fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
bool failed();
void clear_error();
bool check(hipError_t e, const char *what);

hipError_t launch_prep_soa8(const float *f, int source_n, int sz, int order, float4 *pos4,
                            float *mom, hipStream_t st);
hipError_t launch_prep_hexacell(const float *tab, int stride_floats, int64_t n, float4 *pos4,
                                float *mom, hipStream_t st);
hipError_t launch_mn(int order, int soft, int qsign, const float *sinkpos, float *acc, int m,
                     const float4 *pos4, const float *mom, int n, float e, hipStream_t st);
hipError_t launch_mns(const float *sinkpos, float *acc, int m, const uint32_t *idx, int nidx,
                      const float *srcpos, float mmass, float eps_inv, hipStream_t st);
// max_source_n / max_len: the longest source list of any sink (sets the split)
hipError_t launch_mnss(const float *sinkpos, float *acc, int m, const uint16_t *ss_index,
                       const uint32_t *ss_seg, int ss_len, const uint32_t *seg, int seg_len,
                       const float *srcpos, const int *source_n, int64_t max_source_n,
                       float mmass, float eps_inv, int *errflag, hipStream_t st);
hipError_t launch_qnss(int qsign, const float *sinkpos, float *acc, int m, const uint32_t *seg,
                       int64_t max_len, const float4 *pos4, const float *mom, hipStream_t st);
hipError_t launch_qns(int qsign, const float *sinkpos, float *acc, int m, const int *idx,
                      int source_n, const float4 *pos4, const float *mom, hipStream_t st);

// The batched kernels of the deferred calls: `data` is the uploaded batch (lists
// at the offsets of the calls), sinkpos / acc the whole device body table.
// M x N kinds first prepare every call's sources (block_call: the call of each
// 128-source prep block) into pos4 / mom, then evaluate the tasks.
hipError_t launch_mn_batch(int order, int soft, int qsign, const float *data,
                           const DeferredCall *calls, const int *block_call, int nblocks,
                           const DeferredTask *tasks, int ntasks, float4 *pos4, float *mom,
                           const float *sinkpos, float *acc, hipStream_t st);
hipError_t launch_mns_batch(const uint32_t *data, const DeferredCall *calls,
                            const DeferredTask *tasks, int ntasks, const float *sinkpos,
                            float *acc, hipStream_t st);
hipError_t launch_qns_batch(int qsign, const int *data, const DeferredCall *calls,
                            const DeferredTask *tasks, int ntasks, const float4 *pos4,
                            const float *mom, const float *sinkpos, float *acc, hipStream_t st);

int quad_phi_sign();

}  // namespace pgrav
