// Device view of a prepared force evaluation (see pgrav_batch.hip for the
// layout) and the launchers of the batched kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NXCD 8      // MI355X accelerator complex dies (one L2 each)

namespace pgrav {

enum Variant {
    VAR_REL32C = 0,  // FP32 interactions, group-relative coordinates, compensated sums
    VAR_REL32 = 1,   // FP32, group-relative, plain sums
    VAR_ABS32 = 2,   // FP32, absolute coordinates, plain sums (the reference's arithmetic)
    VAR_F64 = 3,     // FP64 everything, one batched launch
    VAR_PLAIN64 = 4  // FP64, one launch per group (the reference's dispatch)
};

struct BatchView {
    const double4 *bpos = nullptr;
    const double4 *cpos = nullptr;
    const float *mom32 = nullptr;
    const double *mom64 = nullptr;
    const int64_t *gfirst = nullptr;
    const int *gcount = nullptr;
    const int64_t *gslot_off = nullptr;  // bisection groups (merge_groups 5): sinks of group g
    const uint32_t *gslot = nullptr;     // are gslot[gslot_off[g] .. gslot_off[g + 1])
    const double4 *gcen = nullptr;
    const int *lpt = nullptr;   // dispatch order
    const int *xoff = nullptr;  // NXCD + 1 slice offsets into lpt
    const int *order_np = nullptr;  // order for one workgroup per group
    // CSR lists of all four kinds in one index array: the entries of kind k
    // (order-4, order-2, order-0 cells, bodies) for group g are
    // idx[off[k*(ngroups+1)+g] .. off[k*(ngroups+1)+g+1])
    const int64_t *off = nullptr;
    const uint32_t *idx = nullptr;
    const uint8_t *img = nullptr;  // periodic = 2: image of each idx entry
    float4 *out32 = nullptr;
    double4 *out64 = nullptr;
    int *counter = nullptr;
    const float *ewald32 = nullptr;   // Ewald correction table (pgrav_ewald.h), FP32 copy
    const double *ewald64 = nullptr;  // FP64
    int unroll = 4;  // order-0 / body loop unroll of the rel32c kernels: 2, 4 or 8
    int tile = 64;   // list entries per LDS round of the rel32c kernels: 32, 64 or 128 (the
                     // periodic-1 kernels: 64; tiles of 32 and 128 run with the unroll of 4)
    double ewratio = 0.05;  // periodic 1: see pgrav_params.ewald_ratio
    int ngroups = 0;
    int soft = 0;
    double eps = 0;
    double G = 1;
    int periodic = 0;  // 1 nearest image + Ewald table; 2 explicit images (img)
    int nimage = 0;
    double box = 1;
};

hipError_t launch_batch(const BatchView &v, int variant, int persistent_blocks, hipStream_t st);
hipError_t launch_group64(const BatchView &v, int g, hipStream_t st);
hipError_t launch_reset_counter(int *c, hipStream_t st);
hipError_t launch_ewald_table(double L, double *t64, float *t32, hipStream_t st);
// raw moment a (nln/cofm.c:660-720 order): its exponents e[a][q] and the index
// of the moment with e[a][q] - 1 (when e[a][q] > 0)
struct RawIndex {
    signed char e[35][3];
    signed char lower[35][3];
};
RawIndex raw_index();

// The adjoint's buffers (device; sorted body order).  Every sum has one writer.
struct AdjointView {
    const float4 *wv = nullptr;  // weights (w, v) per body
    float4 *gsink = nullptr;     // sink-side position gradient per body
    float4 *gpair = nullptr;     // per cell-list entry: the entry's dL/dX_c share
    float4 *gentry = nullptr;    // per body-list entry: dL/d(source position, mass)
    int64_t ncell_entries = 0;   // cell entries precede the body entries in idx
    // transposed cell lists: work item i is node tcells[i] and entries
    // trange[i] of tlist = (group ref, count | kind << 8, entry position lo, hi)
    const int *tcells = nullptr;
    int ntcells = 0;
    const longlong2 *trange = nullptr;
    const uint4 *tlist = nullptr;
    float *gchunk = nullptr;              // 52 floats per work item
    const int64_t *cchunk_off = nullptr;  // node c: work items [cchunk_off[c], cchunk_off[c+1])
    // transposed body lists by leaf: leaf c's segments lseg[loff[c] .. loff[c+1])
    // = (entry position of body offset u0, u0, u1) -- bodies lfirst[c] + u,
    // u0 <= u < u1, at consecutive positions (relative to ncell_entries);
    // ascending positions
    const int64_t *loff = nullptr;
    const uint4 *lseg = nullptr;
    const int64_t *lfirst = nullptr;
    // body j's gather chunks are gbchunk[pstart[j] .. pstart[j+1]); pbody[t] is
    // chunk t's body
    const int64_t *pstart = nullptr;
    const int *pbody = nullptr;
    float4 *gbchunk = nullptr;
    int64_t nbchunks = 0;
    float4 *gbody = nullptr;
    const double *raw = nullptr;  // 35 raw moments per node
    const float *mat = nullptr;   // 46 x 35: raw moments -> prepared record
    RawIndex rix;
    int64_t nnodes = 0;
    float *gnode = nullptr;  // 40 per node: dL/d(raw), dL/dX
    const int *leaf_of = nullptr, *parent = nullptr;
    int64_t n = 0;
    double4 *out = nullptr;  // dL/dx, dL/dm (times G)
};
hipError_t launch_adjoint(const BatchView &v, const AdjointView &A, hipStream_t st);
// caller order <-> sorted slots (slot i holds caller body perm[i]), pgrav_bodies.hip
hipError_t launch_gather_bodies(const double *pos3, const double *mass, const int64_t *perm,
                                int64_t n, int periodic, double box, const double4 *prev,
                                double4 *out, unsigned long long *bad, hipStream_t st);
hipError_t launch_scatter4(const void *src, int src_fp64, const int64_t *perm, int64_t n, void *dst,
                           int dst_fp64, hipStream_t st);
hipError_t launch_gather4(const void *src, int src_fp64, const int64_t *perm, int64_t n,
                          float4 *dst, hipStream_t st);
// ewald64 non-null: periodic (nearest image + Ewald correction, box edge `box`)
hipError_t launch_direct64(const double4 *bpos, int64_t n, const int64_t *sample, int ns, int soft,
                           double eps, double G, double4 *part, int nsplit, double4 *out,
                           const double *ewald64, double box, hipStream_t st);

}  // namespace pgrav
