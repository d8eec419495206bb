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
//                    over the group's sinks, 8 entries at a time through LDS) is
//                    written once per list entry: gpair (cells: dL/dX_c) and
//                    gentry (bodies: dL/dx_j and dL/dm_j).
//  k_adjoint_cells   one wave per listed cell and chunk of <= 128 listing groups
//                    (the transposed lists): the 43 prepared-record
//                    coefficients' gradient dL/dP (the interaction is linear in
//                    P: the hand-written transpose of the evaluation) plus the
//                    chunk's gpair entries, one output record per chunk.
//  k_adjoint_bodychunks, k_adjoint_bodygather  per body: its gentry entries in
//                    list order, read by leaf segments, in balanced chunks.
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

// dL/dP for one (sink, cell) pair, added to beta: the evaluation in
// cell_interact is linear in the prepared record P; with w, v the sink's
// weights these are the coefficients of P in w.a + v.phi (see pgrav_device.h
// for the formulas).  The order-2 and order-0 coefficients are the leading
// ones of the order-4 form, so one branch-free form serves every list kind:
// k2 (1 for orders 2 and 4, else 0) and k34 (1 for order 4) switch the higher
// orders' terms off.  w = v = 0 gives zero terms.
__device__ __forceinline__ void cell_beta_add(float x, float y, float z, float eps2, float wx_,
                                              float wy_, float wz_, float v, float k2, float k34,
                                              float *beta) {
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
    beta[3] += as;
    const float c2 = k2 * rc;
    const float u2 = k2 * (as * ri2 * ri2 + at * rc);
    const float G2x = u2 * x + c2 * avx, G2y = u2 * y + c2 * avy, G2z = u2 * z + c2 * avz;
    beta[4] += G2x * x;
    beta[5] += G2x * y + G2y * x;
    beta[6] += G2x * z + G2z * x;
    beta[7] += G2y * y;
    beta[8] += G2y * z + G2z * y;
    beta[9] += G2z * z;
    const float c3 = k34 * (rc * rb), c4 = c3 * rq;
    const float ri6 = ri2 * ri2 * ri2;
    const float u3 = k34 * (as * ri6) + at * c3;
    const float u4 = k34 * (as * ri6 * ri2) + at * c4;
    const float G3[3] = {u3 * x + c3 * avx, u3 * y + c3 * avy, u3 * z + c3 * avz};
    const float G4[3] = {u4 * x + c4 * avx, u4 * y + c4 * avy, u4 * z + c4 * avz};
    const float A = x2 - z2, B = y2 - z2;
    const float b3[5] = {A, B, x * y, x * z, y * z};
    const float z3 = 3.0f * z2;
    const float U[7] = {x * (x2 - z3), y * (y2 - z3), y * A, z * (3.0f * x2 - z2), x * B, x * y * z,
                        z * (3.0f * y2 - z2)};
#pragma unroll
    for (int r = 0; r < 3; r++) {
#pragma unroll
        for (int k = 0; k < 5; k++) beta[10 + 5 * r + k] += G3[r] * b3[k];
#pragma unroll
        for (int k = 0; k < 7; k++) beta[25 + 7 * r + k] += G4[r] * U[k];
    }
}

// Sorted-body slot of sink `lane` of group g: a run from gfirst, or for bisection
// groups (body sets) the group's entry in the slot table.  `ref` is what the
// transposed cell lists store for a group: its first sink, or its index when the
// groups are body sets.
__device__ __forceinline__ int64_t sink_slot(const BatchView &V, int64_t ref, int lane) {
    return V.gslot ? (int64_t)V.gslot[V.gslot_off[ref] + lane] : ref + lane;
}

// The list entries' source-side shares, summed over the group's sinks without a
// wave reduction per entry: entry k's per-lane values wait in red[k % 8][lane];
// after every 8 entries (and at a tile's end) lane 8 s + p adds sinks
// 8 p .. 8 p + 7 of entry s (in a fixed order, staggered by p across the LDS
// banks), three DPP steps add the eight partial sums within the lane's row half,
// and lane 8 s writes entry s: out[base + s], s < n.
__device__ __forceinline__ void flush_entries(float4 (*red)[64], int n, float4 *__restrict__ out,
                                              int64_t base) {
    __syncthreads();
    const int lane = threadIdx.x, s = lane >> 3, p = lane & 7;
    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const float4 v = red[s][p * 8 + ((i + p) & 7)];
        ax += v.x;
        ay += v.y;
        az += v.z;
        aw += v.w;
    }
    ax = dpp_add<0x141, 0xf>(dpp_add<0x4e, 0xf>(dpp_add<0xb1, 0xf>(ax)));
    ay = dpp_add<0x141, 0xf>(dpp_add<0x4e, 0xf>(dpp_add<0xb1, 0xf>(ay)));
    az = dpp_add<0x141, 0xf>(dpp_add<0x4e, 0xf>(dpp_add<0xb1, 0xf>(az)));
    aw = dpp_add<0x141, 0xf>(dpp_add<0x4e, 0xf>(dpp_add<0xb1, 0xf>(aw)));
    if (p == 0 && s < n)
        out[base + s] = make_float4(ax, ay, az, aw);
    __syncthreads();
}

// Every buffer arrives as its own __restrict__ argument: the kernel stores
// entry results between its reads of the lists and records, and only pointers
// known to be distinct let the compiler read those through scalar loads.
template <int SOFT>
__global__ __launch_bounds__(64) void k_adjoint(
    BatchView V, AdjointView A, const int *__restrict__ order, const int *__restrict__ gcount,
    const double4 *__restrict__ gcen, const int64_t *__restrict__ gfirst,
    const int64_t *__restrict__ gslot_off, const uint32_t *__restrict__ gslot,
    const int64_t *__restrict__ off, const uint32_t *__restrict__ idx,
    const float *__restrict__ mom, const double4 *__restrict__ bpos,
    const double4 *__restrict__ cpos, const float4 *__restrict__ wv, float4 *__restrict__ gpair,
    float4 *__restrict__ gentry, float4 *__restrict__ gsink) {
    __shared__ float4 lds[64];
    __shared__ float4 red[8][64];
    // XCD-aware order: consecutive workgroups land on different XCDs, so the
    // forward's per-XCD slices of Morton-neighbouring groups keep their cells in
    // one XCD's L2 here too
    const int g = order ? order[blockIdx.x] : (int)blockIdx.x;
    const int lane = threadIdx.x;
    const int cnt = gcount[g];
    const double4 gc = gcen[g];
    const bool live = lane < cnt;
    const int ls = live ? lane : 0;
    const int64_t me = gslot ? (int64_t)gslot[gslot_off[g] + ls] : gfirst[g] + ls;
    const double4 sp = bpos[me];
    const float sx = (float)(sp.x - gc.x), sy = (float)(sp.y - gc.y), sz = (float)(sp.z - gc.z);
    const float4 w = live ? wv[me] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float eps = (float)V.eps, eps2 = eps * eps;
    const float eps_inv = SOFT == SOFT_K1 ? 1.0f / eps : 0.0f;
    const float eps2c = SOFT == SOFT_PLUMMER ? eps2 : 0.0f;
    const int64_t ncell = A.ncell_entries;
    float gx = 0.f, gy = 0.f, gz = 0.f;

    for (int kind = 0; kind < 4; kind++) {
        const int64_t *offk = off + (int64_t)kind * (V.ngroups + 1);
        const int64_t o0 = offk[g], o1 = offk[g + 1];
        for (int64_t t0 = o0; t0 < o1; t0 += 64) {
            const int nt = (int)min((int64_t)64, o1 - t0);
            if (lane < nt) {
                const uint32_t e = idx[t0 + lane];
                const double4 s = kind == 3 ? bpos[e] : cpos[e];
                lds[lane] = make_float4((float)(s.x - gc.x), (float)(s.y - gc.y),
                                        (float)(s.z - gc.z), (float)s.w);
            }
            __syncthreads();
            if (kind < L_BODY_KIND) {
                for (int k = 0; k < nt; k++) {
                    const float4 r = lds[k];
                    // the sink offset, differentiated along this sink's weight w
                    const D1 dx(sx - r.x, w.x), dy(sy - r.y, w.y), dz(sz - r.z, w.z);
                    Acc4T<D1> a = {D1(0.f), D1(0.f), D1(0.f), D1(0.f)};
                    const uint32_t e = idx[t0 + k];
                    if (kind == 0)
                        cell_interact<4, false, 1, D1, float>(mom + (size_t)e * PREC4, D1(r.w), dx, dy, dz, D1(eps2c), a);
                    else if (kind == 1)
                        cell_interact<2, false, 1, D1, float>(mom + (size_t)e * PREC4, D1(r.w), dx, dy, dz, D1(eps2c), a);
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
                    red[k & 7][lane] = make_float4(-px, -py, -pz, 0.f);
                    if ((k & 7) == 7 || k == nt - 1)
                        flush_entries(red, (k & 7) + 1, gpair, t0 + (k & ~7));
                }
            } else {
                // bodies, 8 entries per block with the LDS reads of the block in flight
                // together (the pair is ~30 VALU)
                for (int k0 = 0; k0 < nt; k0 += 8) {
                    const int m = min(8, nt - k0);
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        const float4 r = lds[k0 + i];  // k0 + i < 64: inside the tile buffer
                        if (i < m) {
                            const float X = sx - r.x, Y = sy - r.y, Z = sz - r.z;
                            float px, py, pz, pm;
                            if (SOFT == SOFT_K1) {
                                const D1 dx(X, w.x), dy(Y, w.y), dz(Z, w.z);
                                Acc4T<D1> a = {D1(0.f), D1(0.f), D1(0.f), D1(0.f)};
                                // unit mass: the pair is linear in the source mass
                                k1_interact<D1>(D1(1.f), dx, dy, dz, D1(eps_inv), D1(eps2), a);
                                px = r.w * (a.x.d - w.w * a.x.v);
                                py = r.w * (a.y.d - w.w * a.y.v);
                                pz = r.w * (a.z.d - w.w * a.z.v);
                                pm = w.x * a.x.v + w.y * a.y.v + w.z * a.z.v + w.w * a.w.v;
                            } else {
                                // Plummer (or none): with f = 1/sqrt(r^2 + eps^2), a = -m f^3 x and
                                // phi = -m f, so J w - v a = m f^3 ((3 f^2 x.w + v) x - w) and
                                // d(w.a + v phi)/dm = -(f^3 x.w + v f); the unsoftened self pair is 0
                                const float r2 = fmaf(X, X, fmaf(Y, Y, fmaf(Z, Z, eps2c)));
                                const bool self = SOFT == SOFT_NONE && r2 == 0.f;
                                const float f = self ? 0.f : __builtin_amdgcn_rsqf(r2);
                                const float f2 = f * f, f3 = f2 * f;
                                const float xw = w.x * X + w.y * Y + w.z * Z;
                                const float t = 3.0f * f2 * xw + w.w;
                                const float mf3 = r.w * f3;
                                px = mf3 * (t * X - w.x);
                                py = mf3 * (t * Y - w.y);
                                pz = mf3 * (t * Z - w.z);
                                pm = -(f3 * xw + w.w * f);
                            }
                            gx += px;
                            gy += py;
                            gz += pz;
                            red[i][lane] = make_float4(-px, -py, -pz, pm);
                        }
                    }
                    flush_entries(red, m, gentry, t0 + k0 - ncell);
                }
            }
            __syncthreads();
        }
    }
    if (live)
        gsink[me] = make_float4(gx, gy, gz, 0.f);
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
// BODYSET: the groups are body sets (slot table), known at compile time so that
// the run case has no dependent load in its sink prefetch
template <int SOFT, bool BODYSET>
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
        // entry k's sinks (positions, weights), loaded while the entry before it
        // is evaluated: two register sets used in turn, so the loads of one are
        // in flight while the other is read and no copy waits for them
        // (every lane loads, past the last entry too, from a valid slot, so the
        // number of loads in flight is the same on every path)
        struct In {
            double4 sp;
            float4 w;
            float on;
            int kind;
        };
        auto fetch = [&](int k, In &s) {
            uint32_t f;
            int cnt;
            entry(min(k, nt - 1), f, cnt, s.kind);
            if (k >= nt)
                cnt = 0;
            const bool on = lane < cnt;
            const int ls = on ? lane : 0;
            const int64_t me = BODYSET ? (int64_t)V.gslot[V.gslot_off[f] + ls] : (int64_t)f + ls;
            s.sp = V.bpos[me];
            s.w = A.wv[me];
            // lanes without a sink (and entries past the last) add zero terms: the
            // weights are scaled by 0 at use (a product, not a select, so that the
            // load stays unconditional)
            s.on = on ? 1.f : 0.f;
        };
        auto eval = [&](const In &s) {
            const float x = (float)(s.sp.x - X.x), y = (float)(s.sp.y - X.y),
                        z = (float)(s.sp.z - X.z);
            const float k2 = s.kind <= 1 ? 1.f : 0.f, k34 = s.kind == 0 ? 1.f : 0.f;
            cell_beta_add(x, y, z, eps2c, s.on * s.w.x, s.on * s.w.y, s.on * s.w.z, s.on * s.w.w, k2,
                          k34, acc);
        };
        // (the scheduling barriers keep the compiler from moving a fetch above
        // the evaluation it is meant to overlap)
        In ea, eb;
        fetch(0, ea);
        for (int k = 0; k < nt; k += 2) {
            fetch(k + 1, eb);
            __builtin_amdgcn_sched_barrier(0);
            eval(ea);
            __builtin_amdgcn_sched_barrier(0);
            fetch(k + 2, ea);
            __builtin_amdgcn_sched_barrier(0);
            eval(eb);
            __builtin_amdgcn_sched_barrier(0);
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
        out[0] = out[1] = out[2] = 0.f;  // m and the dipole are not in the record's gradient
        out[46] = sx;
        out[47] = sy;
        out[48] = sz;
        out[49] = out[50] = out[51] = 0.f;
    }
}

// Per body: its body-list entries' source gradients, in list order, in two
// passes.  Entry counts per body are very uneven (a few bodies are listed by
// thousands of groups), so a body's leaf segments are cut into at most 64
// chunks of >= 16 segments, one thread each (k_adjoint_bodychunks), and a
// second pass adds each body's chunk sums in order (k_adjoint_bodygather).
// The bodies of one leaf walk the leaf's segments together: one index record
// per segment for all of them and consecutive entries.
__device__ __forceinline__ int64_t body_chunk_len(int64_t nseg) {
    return max((int64_t)16, (nseg + 63) / 64);
}

__global__ void k_adjoint_bodychunks(AdjointView A) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nbchunks)
        return;
    const int j = A.pbody[t];
    const int c = A.leaf_of[j];
    const uint32_t u = (uint32_t)(j - A.lfirst[c]);
    const int64_t q0 = A.loff[c], len = body_chunk_len(A.loff[c + 1] - q0);
    const int64_t s0 = q0 + (t - A.pstart[j]) * len, s1 = min(A.loff[c + 1], s0 + len);
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    // 8 segments at a time: their index records, then their entries, in flight
    // together; summed in list order
    for (int64_t q = s0; q < s1; q += 8) {
        float4 g[8];
        bool in[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint4 sg = q + i < s1 ? A.lseg[q + i] : make_uint4(0u, 1u, 0u, 0u);
            in[i] = u >= sg.y && u < sg.z;
            g[i] = in[i] ? A.gentry[(int64_t)sg.x + (u - sg.y)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (in[i]) {
                s.x += g[i].x;
                s.y += g[i].y;
                s.z += g[i].z;
                s.w += g[i].w;
            }
    }
    A.gbchunk[t] = s;
}

__global__ void k_adjoint_bodygather(AdjointView A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n)
        return;
    const int64_t t1 = A.pstart[j + 1];
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t t0 = A.pstart[j]; t0 < t1; t0 += 8) {
        float4 g[8];
#pragma unroll
        for (int i = 0; i < 8; i++)
            g[i] = t0 + i < t1 ? A.gbchunk[t0 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (t0 + i < t1) {
                s.x += g[i].x;
                s.y += g[i].y;
                s.z += g[i].z;
                s.w += g[i].w;
            }
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
    hipLaunchKernelGGL(k_adjoint<S>, dim3(v.ngroups), dim3(64), 0, st, v, A, v.order_np, v.gcount, \
                       v.gcen, v.gfirst, v.gslot_off, v.gslot, v.off, v.idx, v.mom32, v.bpos,     \
                       v.cpos, A.wv, A.gpair, A.gentry, A.gsink);                                 \
    if (A.ntcells > 0 && v.gslot)                                                                  \
        hipLaunchKernelGGL((k_adjoint_cells<S, true>), dim3(A.ntcells), dim3(64), 0, st, v, A);    \
    else if (A.ntcells > 0)                                                                        \
        hipLaunchKernelGGL((k_adjoint_cells<S, false>), dim3(A.ntcells), dim3(64), 0, st, v, A);
    if (v.soft == SOFT_K1) {
        ADJ(SOFT_K1)
    } else if (v.soft == SOFT_PLUMMER) {
        ADJ(SOFT_PLUMMER)
    } else {
        ADJ(SOFT_NONE)
    }
#undef ADJ
    if (A.nbchunks > 0)
        hipLaunchKernelGGL(k_adjoint_bodychunks, dim3((unsigned)((A.nbchunks + 127) / 128)), dim3(128), 0,
                           st, A);
    hipLaunchKernelGGL(k_adjoint_bodygather, dim3((unsigned)((A.n + 127) / 128)), dim3(128), 0, st, A);
    hipLaunchKernelGGL(k_adjoint_nodes, dim3((unsigned)((A.nnodes + 127) / 128)), dim3(128), 0, st, A);
    hipLaunchKernelGGL(k_adjoint_bodies, dim3((unsigned)((A.n + 127) / 128)), dim3(128), 0, st, v, A);
    return hipGetLastError();
}

}  // namespace pgrav
