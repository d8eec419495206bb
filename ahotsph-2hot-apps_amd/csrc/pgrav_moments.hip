// Upward pass on the GPU, in FP64: centre of mass, raw moments to order 4 about
// it and bmax of every node, from the Morton-sorted bodies, then the prepared
// records the force kernels read.  This is the moment work of nln/cofm.c
// (mpole_add_mono :650-768 for bodies, mpole_add :770-964 for daughters,
// CofmFromDaugh :96-231 for the centre and the daughters' bmax bound), run when
// bodies move under a fixed tree (pgrav_plan_update).
//
// Raw moments are the 35 monomials x^a y^b z^c (a+b+c <= 4) summed with the
// body masses, in the order of nln/cofm.c:660-720.  A daughter's moments about
// its centre c' move to the parent's centre c with the multinomial shift
//   M_alpha(c) = sum_{beta <= alpha} C(alpha, beta) d^(alpha - beta) M_beta(c'),
//   d = c' - c,  C(alpha, beta) = prod_i binom(alpha_i, beta_i),
// exact in exact arithmetic; the tables below are built at compile time.
//
// One thread per node; the nodes of one tree level per launch, deepest level
// first, so every daughter is final before its parent reads it.
#include <hip/hip_runtime.h>

#include "pgrav_moments.h"
#include "pgrav_record.h"

namespace pgrav {

namespace {

struct MomentIndex {
    int e[35][3];
};

constexpr MomentIndex make_index() {
    MomentIndex m{};
    int k = 0;
    for (int order = 0; order <= 4; order++)
        for (int a = order; a >= 0; a--)
            for (int b = order - a; b >= 0; b--) {
                m.e[k][0] = a;
                m.e[k][1] = b;
                m.e[k][2] = order - a - b;
                k++;
            }
    return m;
}

constexpr MomentIndex MI = make_index();

constexpr int index_of(int a, int b, int c) {
    for (int k = 0; k < 35; k++)
        if (MI.e[k][0] == a && MI.e[k][1] == b && MI.e[k][2] == c)
            return k;
    return -1;
}

constexpr double binom(int n, int r) {
    double c = 1;
    for (int i = 0; i < r; i++) c = c * (n - i) / (i + 1);
    return c;
}

// the 35 monomials of (x, y, z)
__device__ __forceinline__ void monomials(double x, double y, double z, double *mono) {
    double px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
    for (int k = 1; k < 5; k++) {
        px[k] = px[k - 1] * x;
        py[k] = py[k - 1] * y;
        pz[k] = pz[k - 1] * z;
    }
#pragma unroll
    for (int k = 0; k < 35; k++) mono[k] = px[MI.e[k][0]] * py[MI.e[k][1]] * pz[MI.e[k][2]];
}

// The 210 (alpha, beta) pairs with beta <= alpha, their multinomial coefficient
// and the index of the monomial d^(alpha - beta).
struct ShiftTerm {
    int al, be, g;
    double c;
};
struct ShiftTable {
    ShiftTerm t[210];
    int n;
};

constexpr ShiftTable make_shift() {
    ShiftTable s{};
    s.n = 0;
    for (int al = 0; al < 35; al++)
        for (int be = 0; be < 35; be++)
            if (MI.e[be][0] <= MI.e[al][0] && MI.e[be][1] <= MI.e[al][1] && MI.e[be][2] <= MI.e[al][2]) {
                s.t[s.n].al = al;
                s.t[s.n].be = be;
                s.t[s.n].g = index_of(MI.e[al][0] - MI.e[be][0], MI.e[al][1] - MI.e[be][1],
                                      MI.e[al][2] - MI.e[be][2]);
                s.t[s.n].c = binom(MI.e[al][0], MI.e[be][0]) * binom(MI.e[al][1], MI.e[be][1]) *
                             binom(MI.e[al][2], MI.e[be][2]);
                s.n++;
            }
    return s;
}

constexpr ShiftTable ST = make_shift();
static_assert(ST.n == 210, "raw moments to order 4: 210 shift terms");

// dst += src (moments about c') shifted to c; dm = monomials of d = c' - c.  Each
// alpha sums its terms in increasing beta order.
__device__ __forceinline__ void shift_add(const double *src, const double *dm, double *dst) {
#pragma unroll
    for (int k = 0; k < 210; k++) dst[ST.t[k].al] += ST.t[k].c * dm[ST.t[k].g] * src[ST.t[k].be];
}

__global__ __launch_bounds__(128) void k_level_moments(MomentsView T, int lo, int hi) {
    const int k = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= hi)
        return;
    const int nd = T.lev_nodes[k];
    double raw[35];
#pragma unroll
    for (int a = 0; a < 35; a++) raw[a] = 0.0;
    double c[3] = {0.0, 0.0, 0.0}, M = 0.0, bm = 0.0;
    if (T.child[nd] < 0) {
        // P2M (nln/cofm.c mpole_add_mono): centre of mass, then moments about it
        const int64_t f = T.first[nd];
        const int n = T.count[nd];
        for (int64_t i = f; i < f + n; i++) {
            const double4 b = T.bpos[i];
            M += b.w;
            c[0] += b.w * b.x;
            c[1] += b.w * b.y;
            c[2] += b.w * b.z;
        }
        for (int q = 0; q < 3; q++) c[q] /= M;
        for (int64_t i = f; i < f + n; i++) {
            const double4 b = T.bpos[i];
            const double x = b.x - c[0], y = b.y - c[1], z = b.z - c[2];
            double mono[35];
            monomials(x, y, z, mono);
#pragma unroll
            for (int a = 0; a < 35; a++) raw[a] += b.w * mono[a];
            bm = fmax(bm, sqrt(x * x + y * y + z * z));
        }
    } else {
        // M2M (nln/cofm.c mpole_add, CofmFromDaugh): daughters' centres and moments
        const int c0 = T.child[nd], nc = T.nchild[nd];
        for (int j = 0; j < nc; j++) {
            const double4 cj = T.cpos[c0 + j];
            M += cj.w;
            c[0] += cj.w * cj.x;
            c[1] += cj.w * cj.y;
            c[2] += cj.w * cj.z;
        }
        for (int q = 0; q < 3; q++) c[q] /= M;
        for (int j = 0; j < nc; j++) {
            const int ch = c0 + j;
            const double4 cj = T.cpos[ch];
            const double d0 = cj.x - c[0], d1 = cj.y - c[1], d2 = cj.z - c[2];
            double dm[35], src[35];
            monomials(d0, d1, d2, dm);
#pragma unroll
            for (int a = 0; a < 35; a++) src[a] = T.raw[(size_t)ch * 35 + a];
            shift_add(src, dm, raw);
            bm = fmax(bm, sqrt(d0 * d0 + d1 * d1 + d2 * d2) + T.bmax[ch]);
        }
    }
    if (T.count[nd] == 1)
        bm = 0.0;
#pragma unroll
    for (int a = 0; a < 35; a++) T.raw[(size_t)nd * 35 + a] = raw[a];
    T.bmax[nd] = bm;
    T.cpos[nd] = make_double4(c[0], c[1], c[2], raw[0]);
}

// prepared records from the raw moments (the same templates the host uses)
__global__ __launch_bounds__(128) void k_prepare_records(MomentsView T, int64_t nn, int order) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn)
        return;
    double raw[35], q[29], out[PREC4];
    for (int a = 0; a < 35; a++) raw[a] = T.raw[(size_t)i * 35 + a];
    const double4 cp = T.cpos[i];
    const double c[3] = {cp.x, cp.y, cp.z};
    // unnormalised record (R = 1): the prepared values do not depend on R
    raw_to_record<double>(raw, c, 1.0, q);
    q[5] = q[6] = q[7] = 0.0;  // the dipole about the centre of mass is zero
    prep_record<double>(q, order, 0, out);
    out[PREC_TRACE] = raw_trace<double>(raw);
    for (int k = 0; k < PREC4; k++) T.mom32[(size_t)i * PREC4 + k] = (float)out[k];
    if (T.mom64)
        for (int k = 0; k < PREC4; k++) T.mom64[(size_t)i * PREC4 + k] = out[k];
}

}  // namespace

hipError_t launch_moments(const MomentsView &T, const int *lev_off, int nlev, int64_t nnodes,
                          int order, hipStream_t st) {
    // lev_off[L] .. lev_off[L+1] are the nodes of level L; deepest first
    for (int L = nlev - 1; L >= 0; L--) {
        const int lo = lev_off[L], hi = lev_off[L + 1];
        if (hi <= lo)
            continue;
        hipLaunchKernelGGL(k_level_moments, dim3((hi - lo + 127) / 128), dim3(128), 0, st, T, lo, hi);
    }
    if (nnodes > 0)
        hipLaunchKernelGGL(k_prepare_records, dim3((unsigned)((nnodes + 127) / 128)), dim3(128), 0, st,
                           T, nnodes, order);
    return hipGetLastError();
}

}  // namespace pgrav
