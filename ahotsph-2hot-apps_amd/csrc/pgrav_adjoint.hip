// Adjoint of the batched evaluation, tree and lists held fixed.
//
// Loss L = sum_i w_i . a_i + v_i phi_i for per-body weights (w_i, v_i).  With the
// lists fixed, a_i and phi_i depend on the bodies through (1) the sink position
// x_i, (2) each body-list source position x_j and mass m_j, and (3) each
// accepted cell's centre of mass X_c and raw moments M_c, which are functions
// of the cell's bodies.
//
// Every sum has one writer and a fixed order -- no atomics -- so two calls give
// bit-identical gradients:
//
//  k_adjoint         one wave per sink group (lane = sink), the forward's tiling
//                    and XCD-aware group order.  Per pair one forward-mode
//                    derivative along the sink's weight w through the very same
//                    cell_interact / mono_interact code gives the sink's
//                    gradient; the pair's source-side share (minus that, summed
//                    over the group's sinks by a DPP wave reduction) is written
//                    once per list entry: gpair (cells: dL/dX_c) and gentry
//                    (bodies: dL/dx_j and dL/dm_j).
//  k_adjoint_cells   one wave per listed cell and chunk of <= 128 listing groups
//                    (the transposed lists): the 43 prepared-record
//                    coefficients' gradient dL/dP (the interaction is linear in
//                    P: the hand-written transpose of the evaluation) plus the
//                    chunk's gpair entries, one output record per chunk.
//  k_adjoint_bodygather  per body: its gentry entries in list order.
//  k_adjoint_nodes   per cell: its chunks in order; dL/dP -> dL/d(raw moments)
//                    through the fixed linear map raw -> record -> prepared
//                    record, and the total dL/dX_c including the centre
//                    dependence of the moments.
//  k_adjoint_bodies  per body: walk up its ancestors, chaining dL/dM_c and dL/dX_c
//                    to dL/dx_j and dL/dm_j (M_c[a] = sum_j m_j (x_j - X_c)^a,
//                    X_c = sum_j m_j x_j / sum_j m_j).
#include "pgrav_batch.h"
#include "pgrav_device.h"

namespace pgrav {

// ---------------------------------------------------------------- dual numbers
// The sink-position gradient of one pair's w.a + v phi needs J^T w + v grad(phi),
// J the Jacobian of a.  Every interaction here is a conservative field, a = -grad(phi)
// (the order-2/4 expansions with the physical quadrupole sign, the monopole, Plummer
// and K1 bodies), so J is the negated Hessian of phi, symmetric: J^T w = J w, the
// derivative of a along w, and grad(phi) = -a.  One forward-mode derivative along w
// therefore gives the whole gradient:  J w - v a.
struct D1 {
    float v, d;  // value and derivative along the seeded direction
    __device__ D1() = default;
    __device__ D1(float a) : v(a), d(0.f) {}
    __device__ D1(float a, float b) : v(a), d(b) {}
};
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ D1 operator-(D1 a) { return D1(-a.v, -a.d); }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ D1 operator*(float s, D1 b) { return D1(s * b.v, s * b.d); }
__device__ __forceinline__ D1 operator*(D1 b, float s) { return s * b; }
__device__ __forceinline__ D1 fma(D1 a, D1 b, D1 c) { return a * b + c; }
__device__ __forceinline__ D1 &operator+=(D1 &a, D1 b) { return a = a + b; }
__device__ __forceinline__ D1 &operator-=(D1 &a, D1 b) { return a = a - b; }
__device__ __forceinline__ bool operator>(D1 a, D1 b) { return a.v > b.v; }
__device__ __forceinline__ bool operator==(D1 a, D1 b) { return a.v == b.v; }
// -1/sqrt(s): derivative 0.5 s^-3/2 = -0.5 f^3
__device__ __forceinline__ D1 nrsq(D1 s) {
    const float f = nrsq(s.v);
    return D1(f, -0.5f * f * f * f * s.d);
}

#define ADJ_CELL 52  // per work item / node: 46 dL/dP, 3 dL/dX (pair part), pad
#define L_BODY_KIND 3  // list kinds 0..2 are cells (orders 4, 2, 0), 3 bodies

// dL/dP for one (sink, cell) pair: the evaluation in cell_interact is linear in
// the prepared record P; with w, v the sink's weights these are the
// coefficients of P in w.a + v.phi (see pgrav_device.h for the formulas).
template <int ORDER>
__device__ __forceinline__ void cell_beta(float x, float y, float z, float eps2, float wx_,
                                          float wy_, float wz_, float v, float *beta) {
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float r2 = x2 + y2 + z2 + eps2;
    const float rinv = nrsq(r2);
    const float ri2 = rinv * rinv;
    const float ra = rinv * ri2;
    const float wdx = wx_ * x + wy_ * y + wz_ * z;
    const float as = ra * wdx + v * rinv;  // dL/dS_phi
    const float rc = 2.0f * ri2, rb = 1.5f * ri2, rq = (4.0f / 3.0f) * ri2;
    const float at = ra * wdx * rc;        // dL/dT'
    const float avx = -ra * wx_, avy = -ra * wy_, avz = -ra * wz_;  // dL/dV
#pragma unroll
    for (int k = 0; k < 46; k++) beta[k] = 0.0f;
    beta[3] = as;
    if (ORDER == 0)
        return;
    const float c2 = rc;
    const float u2 = as * ri2 * ri2 + at * c2;
    const float G2x = u2 * x + c2 * avx, G2y = u2 * y + c2 * avy, G2z = u2 * z + c2 * avz;
    beta[4] = G2x * x;
    beta[5] = G2x * y + G2y * x;
    beta[6] = G2x * z + G2z * x;
    beta[7] = G2y * y;
    beta[8] = G2y * z + G2z * y;
    beta[9] = G2z * z;
    if (ORDER == 2)
        return;
    const float c3 = rc * rb, c4 = rc * rb * rq;
    const float ri6 = ri2 * ri2 * ri2;
    const float u3 = as * ri6 + at * c3;
    const float u4 = as * ri6 * ri2 + at * c4;
    const float G3[3] = {u3 * x + c3 * avx, u3 * y + c3 * avy, u3 * z + c3 * avz};
    const float G4[3] = {u4 * x + c4 * avx, u4 * y + c4 * avy, u4 * z + c4 * avz};
    const float A = x2 - z2, B = y2 - z2;
    const float b3[5] = {A, B, x * y, x * z, y * z};
    const float z3 = 3.0f * z2;
    const float U[7] = {x * (x2 - z3), y * (y2 - z3), y * A, z * (3.0f * x2 - z2), x * B, x * y * z,
                        z * (3.0f * y2 - z2)};
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < 5; k++) beta[10 + 5 * r + k] = G3[r] * b3[k];
        for (int k = 0; k < 7; k++) beta[25 + 7 * r + k] = G4[r] * U[k];
    }
}

// Sorted-body slot of sink `lane` of group g: a run from gfirst, or for bisection
// groups (body sets) the group's entry in the slot table.  `ref` is what the
// transposed cell lists store for a group: its first sink, or its index when the
// groups are body sets.
__device__ __forceinline__ int64_t sink_slot(const BatchView &V, int64_t ref, int lane) {
    return V.gslot ? (int64_t)V.gslot[V.gslot_off[ref] + lane] : ref + lane;
}

template <int SOFT>
__global__ __launch_bounds__(64) void k_adjoint(BatchView V, AdjointView A) {
    __shared__ float4 lds[64];
    // XCD-aware order: consecutive workgroups land on different XCDs, so the
    // forward's per-XCD slices of Morton-neighbouring groups keep their cells in
    // one XCD's L2 here too
    const int g = V.order_np ? V.order_np[blockIdx.x] : (int)blockIdx.x;
    const int lane = threadIdx.x;
    const int cnt = V.gcount[g];
    const double4 gc = V.gcen[g];
    const bool live = lane < cnt;
    const int64_t me = sink_slot(V, V.gslot ? g : V.gfirst[g], live ? lane : 0);
    const double4 sp = V.bpos[me];
    const float sx = (float)(sp.x - gc.x), sy = (float)(sp.y - gc.y), sz = (float)(sp.z - gc.z);
    const float4 w = live ? A.wv[me] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float eps = (float)V.eps, eps2 = eps * eps;
    const float eps_inv = SOFT == SOFT_K1 ? 1.0f / eps : 0.0f;
    const float eps2c = SOFT == SOFT_PLUMMER ? eps2 : 0.0f;
    float gx = 0.f, gy = 0.f, gz = 0.f;

    for (int kind = 0; kind < 4; kind++) {
        const int64_t *offk = V.off + (int64_t)kind * (V.ngroups + 1);
        const int64_t o0 = offk[g], o1 = offk[g + 1];
        const uint32_t *__restrict__ idx = V.idx;
        for (int64_t t0 = o0; t0 < o1; t0 += 64) {
            const int nt = (int)min((int64_t)64, o1 - t0);
            if (lane < nt) {
                const uint32_t e = idx[t0 + lane];
                const double4 s = kind == 3 ? V.bpos[e] : V.cpos[e];
                lds[lane] = make_float4((float)(s.x - gc.x), (float)(s.y - gc.y),
                                        (float)(s.z - gc.z), (float)s.w);
            }
            __syncthreads();
            for (int k = 0; k < nt; k++) {
                const float4 r = lds[k];
                const float X = sx - r.x, Y = sy - r.y, Z = sz - r.z;
                // the sink offset, differentiated along this sink's weight w
                const D1 dx(X, w.x), dy(Y, w.y), dz(Z, w.z);
                Acc4T<D1> a = {D1(0.f), D1(0.f), D1(0.f), D1(0.f)};
                if (kind < L_BODY_KIND) {
                    const uint32_t e = idx[t0 + k];
                    if (kind == 0)
                        cell_interact<4, false, 1, D1, float>(V.mom32 + (size_t)e * PREC4, D1(r.w), dx, dy, dz, D1(eps2c), a);
                    else if (kind == 1)
                        cell_interact<2, false, 1, D1, float>(V.mom32 + (size_t)e * PREC4, D1(r.w), dx, dy, dz, D1(eps2c), a);
                    else
                        mono_interact<D1>(D1(r.w), dx, dy, dz, D1(eps2c), a);
                    // d(w.a + v phi)/d(sink position) = J w - v a; the cell's centre
                    // takes minus that, summed over the group's sinks
                    const float px = a.x.d - w.w * a.x.v;
                    const float py = a.y.d - w.w * a.y.v;
                    const float pz = a.z.d - w.w * a.z.v;
                    gx += px;
                    gy += py;
                    gz += pz;
                    const float cx = wave_sum(-px), cy = wave_sum(-py), cz = wave_sum(-pz);
                    if (lane == 0)
                        A.gpair[t0 + k] = make_float4(cx, cy, cz, 0.f);
                } else {
                    // bodies at unit mass: the pair is linear in the source mass
                    const bool self = X == 0.f && Y == 0.f && Z == 0.f;
                    if (SOFT == SOFT_K1)
                        k1_interact<D1>(D1(1.f), dx, dy, dz, D1(eps_inv), D1(eps2), a);
                    else if (SOFT == SOFT_PLUMMER)
                        mono_interact<D1>(D1(1.f), dx, dy, dz, D1(eps2), a);
                    else if (!self)
                        mono_interact<D1>(D1(1.f), dx, dy, dz, D1(0.f), a);
                    const float px = r.w * (a.x.d - w.w * a.x.v);
                    const float py = r.w * (a.y.d - w.w * a.y.v);
                    const float pz = r.w * (a.z.d - w.w * a.z.v);
                    const float pm = w.x * a.x.v + w.y * a.y.v + w.z * a.z.v + w.w * a.w.v;
                    gx += px;
                    gy += py;
                    gz += pz;
                    const float bx = wave_sum(-px), by = wave_sum(-py), bz = wave_sum(-pz),
                                bm = wave_sum(pm);
                    if (lane == 0)
                        A.gentry[t0 + k - A.ncell_entries] = make_float4(bx, by, bz, bm);
                }
            }
            __syncthreads();
        }
    }
    if (live)
        A.gsink[me] = make_float4(gx, gy, gz, 0.f);
}

// ---- transposed cell lists (built by the host): for every cell, the groups
// that list it, entry = (first sink -- the group's index for body-set groups --,
// sink count | list kind << 8, the entry's position in the lists), so that a
// cell's record gradient is summed by one wave alone.

// One wave per (cell, chunk of <= 128 listing groups): lanes take a listing
// group's sinks (coalesced), the wave walks the chunk's groups; every lane keeps
// its partial dL/dP in registers across groups and the wave sums them once, at
// the end, into the chunk's own output record.  The chunk's list entries arrive
// 64 at a time in one vector load and are read back with v_readlane; the next
// group's sinks load while the current group is evaluated.
template <int SOFT>
__global__ __launch_bounds__(64) void k_adjoint_cells(BatchView V, AdjointView A) {
    const int item = blockIdx.x;
    const int c = A.tcells[item];
    const int64_t t0 = A.trange[item].x, t1 = A.trange[item].y;
    const int lane = threadIdx.x;
    const double4 X = V.cpos[c];
    const float eps2c = SOFT == SOFT_PLUMMER ? (float)(V.eps * V.eps) : 0.0f;
    float acc[46];
#pragma unroll
    for (int q = 0; q < 46; q++) acc[q] = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f;  // the chunk's gpair entries, one per lane
    for (int64_t tb = t0; tb < t1; tb += 64) {
        const int nt = (int)min((int64_t)64, t1 - tb);
        const uint4 mine = lane < nt ? A.tlist[tb + lane] : make_uint4(0u, 0u, 0u, 0u);
        if (lane < nt) {
            const float4 gp = A.gpair[(int64_t)mine.z | ((int64_t)mine.w << 32)];
            px += gp.x;
            py += gp.y;
            pz += gp.z;
        }
        // entry k: first sink, count and kind, wave-uniform
        auto entry = [&](int k, uint32_t &first, int &cnt, int &kind) {
            first = (uint32_t)__builtin_amdgcn_readlane((int)mine.x, k);
            const uint32_t y = (uint32_t)__builtin_amdgcn_readlane((int)mine.y, k);
            cnt = (int)(y & 255u);
            kind = (int)(y >> 8);
        };
        uint32_t f;
        int cnt, kind;
        entry(0, f, cnt, kind);
        double4 sp = make_double4(0, 0, 0, 0);
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane < cnt) {
            const int64_t me = sink_slot(V, f, lane);
            sp = V.bpos[me];
            w = A.wv[me];
        }
        for (int k = 0; k < nt; k++) {
            uint32_t fn = 0;
            int cn = 0, kn = 0;
            double4 spn = make_double4(0, 0, 0, 0);
            float4 wn = make_float4(0.f, 0.f, 0.f, 0.f);
            if (k + 1 < nt) {
                entry(k + 1, fn, cn, kn);
                if (lane < cn) {
                    const int64_t me = sink_slot(V, fn, lane);
                    spn = V.bpos[me];
                    wn = A.wv[me];
                }
            }
            if (lane < cnt) {
                const float x = (float)(sp.x - X.x), y = (float)(sp.y - X.y), z = (float)(sp.z - X.z);
                float beta[46];
                if (kind == 0)
                    cell_beta<4>(x, y, z, eps2c, w.x, w.y, w.z, w.w, beta);
                else if (kind == 1)
                    cell_beta<2>(x, y, z, eps2c, w.x, w.y, w.z, w.w, beta);
                else
                    cell_beta<0>(x, y, z, eps2c, w.x, w.y, w.z, w.w, beta);
#pragma unroll
                for (int q = 3; q < 46; q++) acc[q] += beta[q];
            }
            cnt = cn;
            kind = kn;
            sp = spn;
            w = wn;
        }
    }
    float *out = A.gchunk + (size_t)item * ADJ_CELL;
#pragma unroll
    for (int q = 3; q < 46; q++) {
        const float s = wave_sum(acc[q]);
        if (lane == 0)
            out[q] = s;
    }
    const float sx = wave_sum(px), sy = wave_sum(py), sz = wave_sum(pz);
    if (lane == 0) {
        out[46] = sx;
        out[47] = sy;
        out[48] = sz;
    }
}

// per body: its body-list entries' source gradients, in list order
__global__ void k_adjoint_bodygather(AdjointView A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n)
        return;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t e = A.boff[j]; e < A.boff[j + 1]; e++) {
        const float4 g = A.gentry[A.blist[e]];
        s.x += g.x;
        s.y += g.y;
        s.z += g.z;
        s.w += g.w;
    }
    A.gbody[j] = s;
}

// per node: its chunks' sums in order, dL/d(raw moments) = Mat^T dL/dP and the
// total dL/dX_c.  Mat (46 x 35) is the fixed linear map raw moments -> prepared
// record (R = 1); A.rix holds each raw moment's exponents and the index of the
// moment one order lower per axis.
__global__ void k_adjoint_nodes(AdjointView A) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.nnodes)
        return;
    float gp[ADJ_CELL];
    for (int q = 3; q < 49; q++) gp[q] = 0.f;
    for (int64_t it = A.cchunk_off[c]; it < A.cchunk_off[c + 1]; it++) {
        const float *p = A.gchunk + (size_t)it * ADJ_CELL;
        for (int q = 3; q < 49; q++) gp[q] += p[q];
    }
    float gr[35];
    for (int a = 0; a < 35; a++) {
        float s = 0.f;
        for (int k = 3; k < 46; k++) s += A.mat[k * 35 + a] * gp[k];
        gr[a] = s;
    }
    // dM_a/dX_q = -e_q M_{a - e_q}
    float gX[3] = {gp[46], gp[47], gp[48]};
    const double *raw = A.raw + (size_t)c * 35;
    for (int a = 1; a < 35; a++)
        for (int q = 0; q < 3; q++) {
            const int e = A.rix.e[a][q];
            if (e)
                gX[q] -= gr[a] * (float)e * (float)raw[A.rix.lower[a][q]];
        }
    float *o = A.gnode + c * 40;
    for (int k = 0; k < 35; k++) o[k] = gr[k];
    o[35] = gX[0];
    o[36] = gX[1];
    o[37] = gX[2];
}

// per body: chain through every ancestor's moments and centre of mass
__global__ void k_adjoint_bodies(BatchView V, AdjointView A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n)
        return;
    const double4 p = V.bpos[j];
    const float4 s = A.gsink[j], b = A.gbody[j];
    float g[3] = {s.x + b.x, s.y + b.y, s.z + b.z};
    float gm = b.w;
    const float mj = (float)p.w;
    for (int c = A.leaf_of[j]; c >= 0; c = A.parent[c]) {
        const double4 C = V.cpos[c];
        const float d[3] = {(float)(p.x - C.x), (float)(p.y - C.y), (float)(p.z - C.z)};
        const float *gn = A.gnode + (size_t)c * 40;
        float pw[3][5];
        for (int q = 0; q < 3; q++) {
            pw[q][0] = 1.f;
            for (int k = 1; k < 5; k++) pw[q][k] = pw[q][k - 1] * d[q];
        }
        int a = 0;
        for (int o = 0; o <= 4; o++)
            for (int ex = o; ex >= 0; ex--)
                for (int ey = o - ex; ey >= 0; ey--, a++) {
                    const int ez = o - ex - ey;
                    const float ga = gn[a];
                    gm += ga * pw[0][ex] * pw[1][ey] * pw[2][ez];
                    if (ex) g[0] += ga * mj * ex * pw[0][ex - 1] * pw[1][ey] * pw[2][ez];
                    if (ey) g[1] += ga * mj * ey * pw[0][ex] * pw[1][ey - 1] * pw[2][ez];
                    if (ez) g[2] += ga * mj * ez * pw[0][ex] * pw[1][ey] * pw[2][ez - 1];
                }
        const float Mc = (float)C.w;
        if (Mc != 0.f) {
            for (int q = 0; q < 3; q++) g[q] += gn[35 + q] * mj / Mc;
            gm += (gn[35] * d[0] + gn[36] * d[1] + gn[37] * d[2]) / Mc;
        }
    }
    const double G = V.G;
    A.out[j] = make_double4(G * g[0], G * g[1], G * g[2], G * gm);
}

RawIndex raw_index() {
    RawIndex r{};
    int e[35][3], a = 0;
    for (int o = 0; o <= 4; o++)
        for (int ex = o; ex >= 0; ex--)
            for (int ey = o - ex; ey >= 0; ey--, a++) {
                e[a][0] = ex;
                e[a][1] = ey;
                e[a][2] = o - ex - ey;
            }
    for (int b = 0; b < 35; b++)
        for (int q = 0; q < 3; q++) {
            r.e[b][q] = (signed char)e[b][q];
            r.lower[b][q] = 0;
            if (!e[b][q])
                continue;
            int f[3] = {e[b][0], e[b][1], e[b][2]};
            f[q]--;
            for (int l = 0; l < 35; l++)
                if (e[l][0] == f[0] && e[l][1] == f[1] && e[l][2] == f[2])
                    r.lower[b][q] = (signed char)l;
        }
    return r;
}

hipError_t launch_adjoint(const BatchView &v, const AdjointView &A, hipStream_t st) {
    if (v.ngroups <= 0)
        return hipSuccess;
#define ADJ(S)                                                                                     \
    hipLaunchKernelGGL(k_adjoint<S>, dim3(v.ngroups), dim3(64), 0, st, v, A);                      \
    if (A.ntcells > 0)                                                                             \
        hipLaunchKernelGGL(k_adjoint_cells<S>, dim3(A.ntcells), dim3(64), 0, st, v, A);
    if (v.soft == SOFT_K1) {
        ADJ(SOFT_K1)
    } else if (v.soft == SOFT_PLUMMER) {
        ADJ(SOFT_PLUMMER)
    } else {
        ADJ(SOFT_NONE)
    }
#undef ADJ
    hipLaunchKernelGGL(k_adjoint_bodygather, dim3((unsigned)((A.n + 127) / 128)), dim3(128), 0, st, A);
    hipLaunchKernelGGL(k_adjoint_nodes, dim3((unsigned)((A.nnodes + 127) / 128)), dim3(128), 0, st, A);
    hipLaunchKernelGGL(k_adjoint_bodies, dim3((unsigned)((A.n + 127) / 128)), dim3(128), 0, st, v, A);
    return hipGetLastError();
}

}  // namespace pgrav
