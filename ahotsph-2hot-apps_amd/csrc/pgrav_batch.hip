// Batched evaluation of every sink group's interaction list in one launch.
//
// Layout in HBM (all device-resident for the whole force evaluation):
//   bpos   double4[N]      Morton-sorted bodies: FP64 x, y, z and mass
//   cpos   double4[Nnode]  cell centre of mass (FP64) and mass
//   mom32  float[Nnode*48] prepared multipole records (pgrav_device.h); mom64 FP64 twin
//   gfirst/gcount/gcen     group body range and FP64 centre + radius (or, for
//                          bisection groups, gslot_off/gslot: the group's body slots)
//   off[k]/idx[k]          CSR lists per group: order-4 cells, order-2 cells,
//                          order-0 cells, bodies (u32 node / sorted-body slots)
//   lpt    int[G]          groups by decreasing cost (longest first)
//
// Execution: one wave64 per group, one lane per sink (groups hold <= 64).
// A list is consumed in tiles of 64 entries: each lane fetches one entry's FP64
// centre, subtracts the FP64 group centre, rounds the small offset to FP32 and
// parks it in LDS; then every lane sweeps the tile, reading the offset as an
// LDS broadcast and the 48-float record through wave-uniform scalar loads.
// Interactions run in FP32 on group-relative coordinates; tile partial sums
// are folded into the result with an error-free 2Sum (compensated FP32).
//
// Variants (template): FP32 or FP64 arithmetic, group-relative or absolute
// coordinates, compensated or plain accumulation, Plummer / K1 / no
// softening, periodic (Ewald table) or isolated, persistent work queue or one
// workgroup per group.
#include <math.h>

#include "pgrav_batch.h"
#include "pgrav_device.h"

namespace pgrav {

template <typename T>
struct V4 {
    T x, y, z, w;
};

struct BatchArgs {
    const double4 *bpos;
    const double4 *cpos;
    const void *mom;  // float* or double* prepared records
    const int64_t *gfirst;
    const int *gcount;
    // bisection groups: the sinks of group g are the slots gslot[gslot_off[g] ..]
    // (null: the run gfirst[g] .. gfirst[g] + gcount[g])
    const int64_t *gslot_off;
    const uint32_t *gslot;
    const double4 *gcen;
    const int *lpt;   // dispatch order of the groups
    const int *xoff;  // NXCD + 1 slice offsets into lpt (persistent launch)
    const int *order_np;  // dispatch order of the one-workgroup-per-group launch
    const int64_t *off;  // [4][ngroups+1], absolute positions in idx
    const uint32_t *idx;
    void *out;  // float4* or double4*, sorted body order
    int *counter;
    const float4 *ewald;  // 64^3 correction table (periodic)
    const double4 *ewald64;  // its FP64 original, read by the FP64 kernels
    const uint8_t *img;  // explicit images: the image of each entry, parallel to idx (or null)
    int img_f;           // 2 nimage + 1 images per side
    int unroll;          // order-0 / body loop unroll of the rel32c kernels: 2, 4 or 8
    int ewgroup;         // periodic 1: cells' Ewald corrections expanded about the group centre
    int ngroups;
    int soft;
    double eps;
    double G;
    double box;
};

// nearest periodic image of a displacement (identity when not periodic)
template <bool PERIODIC, typename T>
__device__ __forceinline__ T pbc(T x, T box, T ibox) {
    if (PERIODIC)
        return x - box * rint(x * ibox);
    return x;
}

// ---- Ewald correction (periodic).  The table holds, at the nodes of a
// EWALD_N^3 grid over the positive octant [0, L/2]^3, the correction
// T = (C(r), psi(r)) -- C = f_ewald(r) - r/|r|^3, psi = phi_ewald(r) + 1/|r| --
// and its three partial derivatives (4 x TV per node, built in FP64 by
// k_ewald_table).  The lookup takes the nearest node and its first-order
// Taylor step: one 64-byte record, 12 FMAs, error O(h^2 |T''|) as trilinear
// interpolation over 8 nodes would have.  x = sink - source, already the
// nearest image (|x_i| <= L/2), so the Newtonian term the caller added and
// this correction refer to the same image.
// The potential also takes the trace part of the second-order step: psi's
// Hessian is C's Jacobian, whose trace is -4 pi / L^3 everywhere off the lattice
// points, so (1/2) d.H.d = -(2 pi / 3) |d|^2 / L^3 plus a traceless rest two
// orders smaller; the forces' second-order term has no trace part (div grad C =
// grad div C = 0).
template <typename T>
__device__ __forceinline__ T psi_trace_coef(T box) {
    return T(-2.0 * M_PI / 3.0) / (box * box * box);
}

template <typename T, typename TV>
__device__ __forceinline__ void ewald_corr(const TV *__restrict__ tab, T m, T x, T y, T z,
                                           T box, Acc4T<T> &a) {
    const T h = T(0.5) * box / T(EWALD_N - 1);
    const T inv_h = T(1) / h;
    const T sx = x < 0 ? T(-1) : T(1), sy = y < 0 ? T(-1) : T(1), sz = z < 0 ? T(-1) : T(1);
    const T ax = fabs(x), ay = fabs(y), az = fabs(z);
    // clamp both ways: a NaN offset must not turn into an out-of-range index
    const int i = max(0, min((int)rint(ax * inv_h), EWALD_N - 1));
    const int j = max(0, min((int)rint(ay * inv_h), EWALD_N - 1));
    const int k = max(0, min((int)rint(az * inv_h), EWALD_N - 1));
    const T du = ax - T(i) * h, dv = ay - T(j) * h, dw = az - T(k) * h;
    const TV *p = tab + 4 * (((size_t)i * EWALD_N + j) * EWALD_N + k);
    const TV t0 = p[0], gu = p[1], gv = p[2], gw = p[3];
    const T cx = T(t0.x) + du * T(gu.x) + dv * T(gv.x) + dw * T(gw.x);
    const T cy = T(t0.y) + du * T(gu.y) + dv * T(gv.y) + dw * T(gw.y);
    const T cz = T(t0.z) + du * T(gu.z) + dv * T(gv.z) + dw * T(gw.z);
    const T cp = T(t0.w) + du * T(gu.w) + dv * T(gv.w) + dw * T(gw.w) +
                 psi_trace_coef(box) * (du * du + dv * dv + dw * dw);
    // C is odd per axis; the pull on the sink is m C(r) at r = source - sink = -x
    a.x -= m * sx * cx;
    a.y -= m * sy * cy;
    a.z -= m * sz * cz;
    a.w += m * cp;
}

// The same lookup at the group centre for a cell list entry, as a first-order
// expansion in the sink's offset s from that centre: out[0] = (C, psi) at x0 =
// centre - source, out[1..3] = row i of the Jacobian d(C_i, psi)/dx:
// (dC_i/dx, dC_i/dy, dC_i/dz, dpsi/dx_i).  With the table's octant symmetry (C_i
// odd in x_i, even in the others; psi even) d(sgn_i C_i)/dx_j = sgn_i sgn_j dC_i/du_j.
// A sink then takes C(x0 + s) = C + J s: 12 FMAs on values the tile loader left in
// LDS, instead of a 64-byte gather per pair.  Cells are accepted only well inside
// half a box, so every sink of the group sees the cell in the same image as the
// centre.
template <typename T, typename TV>
__device__ __forceinline__ void ewald_taylor(const TV *__restrict__ tab, T x, T y, T z, T box,
                                             V4<T> *out) {
    const T h = T(0.5) * box / T(EWALD_N - 1);
    const T inv_h = T(1) / h;
    const T sx = x < 0 ? T(-1) : T(1), sy = y < 0 ? T(-1) : T(1), sz = z < 0 ? T(-1) : T(1);
    const T ax = fabs(x), ay = fabs(y), az = fabs(z);
    const int i = max(0, min((int)rint(ax * inv_h), EWALD_N - 1));
    const int j = max(0, min((int)rint(ay * inv_h), EWALD_N - 1));
    const int k = max(0, min((int)rint(az * inv_h), EWALD_N - 1));
    const T du = ax - T(i) * h, dv = ay - T(j) * h, dw = az - T(k) * h;
    const TV *p = tab + 4 * (((size_t)i * EWALD_N + j) * EWALD_N + k);
    const TV t0 = p[0], gu = p[1], gv = p[2], gw = p[3];
    V4<T> c;
    c.x = sx * (T(t0.x) + du * T(gu.x) + dv * T(gv.x) + dw * T(gw.x));
    c.y = sy * (T(t0.y) + du * T(gu.y) + dv * T(gv.y) + dw * T(gw.y));
    c.z = sz * (T(t0.z) + du * T(gu.z) + dv * T(gv.z) + dw * T(gw.z));
    c.w = T(t0.w) + du * T(gu.w) + dv * T(gv.w) + dw * T(gw.w) +
          psi_trace_coef(box) * (du * du + dv * dv + dw * dw);
    out[0] = c;
    V4<T> r;
    r.x = T(gu.x);
    r.y = sx * sy * T(gv.x);
    r.z = sx * sz * T(gw.x);
    r.w = sx * T(gu.w);
    out[1] = r;
    r.x = sy * sx * T(gu.y);
    r.y = T(gv.y);
    r.z = sy * sz * T(gw.y);
    r.w = sy * T(gv.w);
    out[2] = r;
    r.x = sz * sx * T(gu.z);
    r.y = sz * sy * T(gv.z);
    r.z = T(gw.z);
    r.w = sz * T(gw.w);
    out[3] = r;
}

// A cell's Ewald correction for one sink from the group-centre expansion in LDS:
// the pull is m C at x = sink - source (as ewald_corr applies it).  s2q is the
// sink's own second-order potential term, psi_trace_coef |s|^2.
template <typename T>
__device__ __forceinline__ void ewald_apply(const V4<T> *e, T m, T sx, T sy, T sz, T s2q,
                                            Acc4T<T> &a) {
    const V4<T> c = e[0], jx = e[1], jy = e[2], jz = e[3];
    const T cx = c.x + jx.x * sx + jx.y * sy + jx.z * sz;
    const T cy = c.y + jy.x * sx + jy.y * sy + jy.z * sz;
    const T cz = c.z + jz.x * sx + jz.y * sy + jz.z * sz;
    const T cp = c.w + jx.w * sx + jy.w * sy + jz.w * sz + s2q;
    a.x -= m * cx;
    a.y -= m * cy;
    a.z -= m * cz;
    a.w += m * cp;
}

// the table a kernel reads: FP32 arithmetic its FP32 copy, FP64 the FP64 original
__device__ __forceinline__ const float4 *ewald_table(const BatchArgs &A, float) {
    return A.ewald;
}
__device__ __forceinline__ const double4 *ewald_table(const BatchArgs &A, double) {
    return A.ewald64;
}

template <typename T>
__device__ __forceinline__ void fold(Acc4T<T> &s, Acc4T<T> &c, const Acc4T<T> &p, bool comp) {
    if (comp) {
        two_sum_add(s.x, c.x, p.x);
        two_sum_add(s.y, c.y, p.y);
        two_sum_add(s.z, c.z, p.z);
        two_sum_add(s.w, c.w, p.w);
    } else {
        s.x += p.x;
        s.y += p.y;
        s.z += p.z;
        s.w += p.w;
    }
}

// Group-relative (REL) or absolute offset of a source centre, rounded to T.
template <typename T, bool REL, bool PERIODIC>
__device__ __forceinline__ V4<T> source_offset(double4 c, double4 gc, double box) {
    double dx = c.x, dy = c.y, dz = c.z;
    if (REL || PERIODIC) {
        dx -= gc.x;
        dy -= gc.y;
        dz -= gc.z;
    }
    if (PERIODIC) {
        dx -= box * rint(dx / box);
        dy -= box * rint(dy / box);
        dz -= box * rint(dz / box);
        if (!REL) {  // absolute variant: re-add the group centre after wrapping
            dx += gc.x;
            dy += gc.y;
            dz += gc.z;
        }
    }
    return V4<T>{T(dx), T(dy), T(dz), T(c.w)};
}

// PER: 0 isolated, 1 nearest image + Ewald correction, 2 explicit images
// BU: unroll factor of the order-0 and body loops (2, 4 or 8; the autotuner picks it
// for the group-relative compensated FP32 kernels, the others use 4)
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU>
__device__ __forceinline__ void eval_group(const BatchArgs &A, int g, V4<T> *lds, V4<T> *ewl) {
    constexpr bool PERIODIC = PER == 1;
    // group-relative periodic with A.ewgroup: the cells' Ewald corrections come
    // expanded about the group centre (ewl: 4 V4 per tile entry, then one flag per
    // entry saying the expansion is used and holds for every sink); bodies keep the
    // per-pair lookup
    constexpr bool EW_GROUP = PERIODIC && REL;
    const int lane = threadIdx.x;
    const int cnt = A.gcount[g];
    const double4 gc = A.gcen[g];
    const bool live = lane < cnt;
    int64_t slot;  // this lane's sink
    if (A.gslot)
        slot = A.gslot[A.gslot_off[g] + (live ? lane : 0)];
    else
        slot = A.gfirst[g] + (live ? lane : 0);
    const double4 sp = A.bpos[slot];
    T sx, sy, sz;
    if (REL) {
        sx = T(sp.x - gc.x);
        sy = T(sp.y - gc.y);
        sz = T(sp.z - gc.z);
    } else {
        sx = T(sp.x);
        sy = T(sp.y);
        sz = T(sp.z);
    }
    const T eps = T(A.eps);
    const T eps2 = eps * eps;
    const T eps_inv = SOFT == SOFT_K1 ? T(1) / eps : T(0);
    const T eps2_cell = SOFT == SOFT_PLUMMER ? eps2 : T(0);
    const T box = T(A.box);
    const T ibox = T(1) / box;
    const T *__restrict__ mom = (const T *)A.mom;
    Acc4T<T> s = {0, 0, 0, 0}, c = {0, 0, 0, 0};
    // periodic: this lane's share of the listed cells' second-moment traces (the
    // potential's second-order Ewald term, the same for every sink of the group)
    double trace = 0.0;
    const T s2q = EW_GROUP ? psi_trace_coef(box) * (sx * sx + sy * sy + sz * sz) : T(0);

#pragma unroll 1
    for (int kind = 0; kind < 4; kind++) {
        const int64_t *offk = A.off + (int64_t)kind * (A.ngroups + 1);
        const uint32_t *__restrict__ idx = A.idx;
        const int64_t o0 = offk[g], o1 = offk[g + 1];
        for (int64_t t0 = o0; t0 < o1; t0 += 64) {
            const int nt = (int)min((int64_t)64, o1 - t0);
            if (lane < nt) {
                const uint32_t e = idx[t0 + lane];
                double4 src = kind == 3 ? A.bpos[e] : A.cpos[e];
                if (PER == 2) {
                    // image (i, j, k) in nln/mac.c:91-111 order: the source sits at
                    // x + (i, j, k) box, added in FP64
                    const int im = A.img[t0 + lane], f = A.img_f, h = f >> 1;
                    src.x += (double)(im / (f * f) - h) * A.box;
                    src.y += (double)((im / f) % f - h) * A.box;
                    src.z += (double)(im % f - h) * A.box;
                }
                lds[lane] = source_offset<T, REL, PERIODIC>(src, gc, A.box);
                if (PERIODIC && kind < 3)
                    trace += (double)mom[(size_t)e * PREC4 + PREC_TRACE];
                if (EW_GROUP && kind < 3 && A.ewgroup) {
                    const V4<T> o = lds[lane];
                    const T x0 = pbc<PERIODIC>(-o.x, box, ibox), y0 = pbc<PERIODIC>(-o.y, box, ibox),
                            z0 = pbc<PERIODIC>(-o.z, box, ibox);
                    ewald_taylor<T>(ewald_table(A, T(0)), x0, y0, z0, box, ewl + 4 * lane);
                    // the expansion holds only if no sink of the group sees another image
                    // of the cell: every sink within the group radius of the centre stays
                    // on the centre's side of the half-box boundary in each axis
                    const T lim = T(0.5) * box - T(gc.w);
                    ewl[4 * 64 + lane].x = (fabs(x0) < lim && fabs(y0) < lim && fabs(z0) < lim) ? T(1) : T(0);
                } else if (EW_GROUP) {
                    ewl[4 * 64 + lane].x = T(0);  // per-pair lookups for this tile
                }
            }
            __syncthreads();
            Acc4T<T> p = {0, 0, 0, 0};
            if (kind == 0) {
                for (int k = 0; k < nt; k++) {
                    const uint32_t e = idx[t0 + k];
                    const V4<T> r = lds[k];
                    const T x = pbc<PERIODIC>(sx - r.x, box, ibox);
                    const T y = pbc<PERIODIC>(sy - r.y, box, ibox);
                    const T z = pbc<PERIODIC>(sz - r.z, box, ibox);
                    cell_interact<4, false, 1, T>(mom + (size_t)e * PREC4, r.w, x, y, z, eps2_cell, p);
                    if (EW_GROUP && ewl[4 * 64 + k].x != T(0))
                        ewald_apply<T>(ewl + 4 * k, r.w, sx, sy, sz, s2q, p);
                    else if (PERIODIC)
                        ewald_corr<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else if (kind == 1) {
                for (int k = 0; k < nt; k++) {
                    const uint32_t e = idx[t0 + k];
                    const V4<T> r = lds[k];
                    const T x = pbc<PERIODIC>(sx - r.x, box, ibox);
                    const T y = pbc<PERIODIC>(sy - r.y, box, ibox);
                    const T z = pbc<PERIODIC>(sz - r.z, box, ibox);
                    cell_interact<2, false, 1, T>(mom + (size_t)e * PREC4, r.w, x, y, z, eps2_cell, p);
                    if (EW_GROUP && ewl[4 * 64 + k].x != T(0))
                        ewald_apply<T>(ewl + 4 * k, r.w, sx, sy, sz, s2q, p);
                    else if (PERIODIC)
                        ewald_corr<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else if (kind == 2) {
#pragma unroll BU
                for (int k = 0; k < nt; k++) {
                    const V4<T> r = lds[k];
                    const T x = pbc<PERIODIC>(sx - r.x, box, ibox);
                    const T y = pbc<PERIODIC>(sy - r.y, box, ibox);
                    const T z = pbc<PERIODIC>(sz - r.z, box, ibox);
                    mono_interact<T>(r.w, x, y, z, eps2_cell, p);
                    if (EW_GROUP && ewl[4 * 64 + k].x != T(0))
                        ewald_apply<T>(ewl + 4 * k, r.w, sx, sy, sz, s2q, p);
                    else if (PERIODIC)
                        ewald_corr<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else {
                // bodies: ~18 VALU per pair, so unroll to keep several LDS reads in flight
#pragma unroll BU
                for (int k = 0; k < nt; k++) {
                    const V4<T> r = lds[k];
                    const T x = pbc<PERIODIC>(sx - r.x, box, ibox);
                    const T y = pbc<PERIODIC>(sy - r.y, box, ibox);
                    const T z = pbc<PERIODIC>(sz - r.z, box, ibox);
                    if (SOFT == SOFT_K1) {
                        k1_interact<T>(r.w, x, y, z, eps_inv, eps2, p);
                    } else if (SOFT == SOFT_PLUMMER) {
                        mono_interact<T>(r.w, x, y, z, eps2, p);
                    } else {  // unsoftened: drop the self pair
                        const bool self = x == T(0) && y == T(0) && z == T(0);
                        mono_interact<T>(self ? T(0) : r.w, x, y, z, self ? T(1) : T(0), p);
                    }
                    // the self pair too: at r = 0 the correction is its limit, C = 0 and
                    // psi(0) = 2.8373/L (the table's origin node), the body's potential
                    // from its own images and the background
                    if (PERIODIC)
                        ewald_corr<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            }
            __syncthreads();
            fold<T>(s, c, p, COMP);
        }
    }
    if (PERIODIC) {
        // the cells' corrections were taken at their centres of mass; the second-
        // order remainder of the potential is (1/2) H : M2 summed over the cells,
        // with H the correction's Hessian (trace -4 pi / L^3 everywhere) and M2 the
        // cell's second moments: its trace part, -(2 pi / 3) / L^3 sum m |x - cofm|^2,
        // is added here (the traceless part is O(1e-2) of it)
        for (int k = 32; k >= 1; k >>= 1) trace += __shfl_xor(trace, k, 64);
        const double b3 = A.box * A.box * A.box;
        fold<T>(s, c, Acc4T<T>{T(0), T(0), T(0), T(-2.0 * M_PI / 3.0 / b3 * trace)}, COMP);
    }
    if (live) {
        const T Gn = T(A.G);  // nln/mac.c:269-277: results scaled by G at the end
        V4<T> r = {Gn * (s.x + c.x), Gn * (s.y + c.y), Gn * (s.z + c.z), Gn * (s.w + c.w)};
        ((V4<T> *)A.out)[slot] = r;
    }
}

// Minimum waves per SIMD asked of the register allocator: left to it, the FP32
// periodic kernels keep the Ewald records in flight with ~80 VGPRs (6 waves);
// asked for 8 they fit in 64 (the K1 persistent variant spills 4 registers).
#ifndef PGRAV_MINWAVES_FP32
#define PGRAV_MINWAVES_FP32 1  // the FP32 kernels without the Ewald table
#endif
template <typename T, int PER>
struct MinWaves {
    static constexpr int value = (sizeof(T) == 4 && PER == 1) ? 8 : (sizeof(T) == 4 ? PGRAV_MINWAVES_FP32 : 1);
};

// one workgroup per group, groups taken in longest-first order
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MinWaves<T, PER>::value)))
void k_batch(BatchArgs A) {
    __shared__ V4<T> lds[64];
    __shared__ V4<T> ewl[PER == 1 && REL ? 5 * 64 : 1];
    eval_group<T, REL, COMP, SOFT, PER, BU>(A, A.order_np[blockIdx.x], lds, ewl);
}

// persistent: a fixed grid drains per-XCD work counters.  Workgroup b runs on
// XCD b % NXCD (the dispatcher deals workgroups to the XCDs round-robin) and
// takes groups from that XCD's slice [xoff[x], xoff[x+1]) of the dispatch order
// first: the host cuts the Morton-ordered groups into NXCD contiguous slices of
// equal estimated cost, so neighbouring groups -- which share most of their
// cells -- run under one XCD's L2.  When its slice is drained it takes the
// remaining groups of the other slices, so an XCD whose estimate was low does
// not idle through the tail.  The last workgroup to leave resets the counters so
// the launch can be replayed (e.g. from a graph).
//
// Every buffer also arrives as its own __restrict__ argument: the loop stores
// a group's results and bumps the counters before it reads the next group's
// lists and records, and only pointers known to be distinct let the compiler
// keep those reads on wave-uniform scalar loads.
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MinWaves<T, PER>::value)))
void k_batch_persistent(
    BatchArgs A, const double4 *__restrict__ bpos, const double4 *__restrict__ cpos,
    const T *__restrict__ mom, const int64_t *__restrict__ gfirst, const int *__restrict__ gcount,
    const double4 *__restrict__ gcen, const int *__restrict__ lpt, const int *__restrict__ xoff,
    const int64_t *__restrict__ off, const uint32_t *__restrict__ idx, V4<T> *__restrict__ out,
    int *__restrict__ counter, const float4 *__restrict__ ewald) {
    A.bpos = bpos;
    A.cpos = cpos;
    A.mom = mom;
    A.gfirst = gfirst;
    A.gcount = gcount;
    A.gcen = gcen;
    A.lpt = lpt;
    A.xoff = xoff;
    A.off = off;
    A.idx = idx;
    A.out = out;
    A.counter = counter;
    A.ewald = ewald;
    __shared__ V4<T> lds[64];
    __shared__ V4<T> ewl[PER == 1 && REL ? 5 * 64 : 1];
    __shared__ int w;
    const int x = blockIdx.x % NXCD;
    // own slice first; once it is drained, the next group of the other XCDs' slices
    // in turn (each counter only grows, so a drained slice stays drained)
    int from = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            int it = -1;
            for (; from < NXCD; from++) {
                const int xx = (x + from) % NXCD;
                const int v = A.xoff[xx] + atomicAdd(&A.counter[2 + xx], 1);
                if (v < A.xoff[xx + 1]) {
                    it = v;
                    break;
                }
            }
            w = it;
        }
        __syncthreads();
        // the LDS value is the same in every lane: say so, so that the group's
        // scalars (list offsets, record addresses) stay in SGPRs and the
        // records keep coming through scalar loads
        const int item = __builtin_amdgcn_readfirstlane(w);
        __syncthreads();
        if (item < 0)
            break;
        eval_group<T, REL, COMP, SOFT, PER, BU>(A, A.lpt[item], lds, ewl);
    }
    if (threadIdx.x == 0) {
        const int done = atomicAdd(&A.counter[1], 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(&A.counter[1], 0);
            for (int k = 0; k < NXCD; k++) atomicExch(&A.counter[2 + k], 0);
        }
    }
}

// the FP64 "plain" reference dispatch: one launch per group (nln/pgrav.cu:1078)
template <int SOFT, int PER>
__global__ __launch_bounds__(64) void k_group64(BatchArgs A, int g) {
    __shared__ V4<double> lds[64];
    eval_group<double, false, false, SOFT, PER, 4>(A, g, lds, nullptr);
}

__global__ void k_reset_counter(int *c) {
    if (threadIdx.x < 2 + NXCD)
        c[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- FP64 direct sum
// Sampled FP64 direct sum (pattern of nln/grav_n2.c:162-173): grid.y splits the
// sources; partial sums go to part[split][sink] and are reduced in fixed order.
// Periodic: nearest image of every source plus the Ewald correction from the
// FP64 table (the tree kernels read its FP32 copy), so the difference to the
// tree result is the tree's own approximation error.
template <int SOFT, bool PERIODIC>
__global__ __launch_bounds__(64) void k_direct64(const double4 *__restrict__ bpos, int64_t n,
                                                 const int64_t *__restrict__ sample, int ns,
                                                 double eps, double4 *__restrict__ part,
                                                 int64_t chunk, const double4 *__restrict__ tab,
                                                 double box) {
    __shared__ double4 lds[64];
    const int s = blockIdx.x * 64 + threadIdx.x;
    const bool live = s < ns;
    const double4 p = bpos[sample[live ? s : 0]];
    const int64_t j0 = blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    const double eps2 = eps * eps, eps_inv = eps > 0 ? 1.0 / eps : 0.0;
    Acc4T<double> a = {0, 0, 0, 0};
    for (int64_t t0 = j0; t0 < j1; t0 += 64) {
        const int nt = (int)min((int64_t)64, j1 - t0);
        if ((int)threadIdx.x < nt)
            lds[threadIdx.x] = bpos[t0 + threadIdx.x];
        __syncthreads();
        for (int k = 0; k < nt; k++) {
            const double4 q = lds[k];
            const double ib = PERIODIC ? 1.0 / box : 0.0;
            const double x = pbc<PERIODIC>(p.x - q.x, box, ib), y = pbc<PERIODIC>(p.y - q.y, box, ib),
                         z = pbc<PERIODIC>(p.z - q.z, box, ib);
            const bool self = x == 0.0 && y == 0.0 && z == 0.0;
            if (SOFT == SOFT_K1) {
                k1_interact<double>(q.w, x, y, z, eps_inv, eps2, a);
            } else if (SOFT == SOFT_PLUMMER) {
                mono_interact<double>(q.w, x, y, z, eps2, a);
            } else {
                mono_interact<double>(self ? 0.0 : q.w, x, y, z, self ? 1.0 : 0.0, a);
            }
            if (PERIODIC)  // the self pair included: C(0) = 0, psi(0) = 2.8373/L
                ewald_corr<double>(tab, q.w, x, y, z, box, a);
        }
        __syncthreads();
    }
    if (live)
        part[(size_t)blockIdx.y * ns + s] = make_double4(a.x, a.y, a.z, a.w);
}

__global__ void k_direct_reduce(const double4 *__restrict__ part, int ns, int nsplit, double G,
                                double4 *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns)
        return;
    double4 a = make_double4(0, 0, 0, 0);
    for (int k = 0; k < nsplit; k++) {
        const double4 q = part[(size_t)k * ns + s];
        a.x += q.x;
        a.y += q.y;
        a.z += q.z;
        a.w += q.w;
    }
    out[s] = make_double4(G * a.x, G * a.y, G * a.z, G * a.w);
}

// ---------------------------------------------------------------- Ewald table
// FP64 restatement of nln/ewald.c:13-64 at one point r: C(r) = f_ewald(r) - r/|r|^3
// and psi(r) = phi_ewald(r) + 1/|r| (at r = 0 the limits 0 and 2 x 1.4186487533 / L).
__device__ double4 ewald_point(double x0, double x1, double x2, double L) {
    const double x[3] = {x0, x1, x2};
    double f[3] = {0, 0, 0}, phi;
    const int NN = 4;
    const double alpha = 2.0 / L;
    phi = M_PI / (alpha * alpha * L * L * L);
    const double r2x = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (r2x == 0.0)
        return make_double4(0, 0, 0, 2.0 * 1.4186487533 / L);
    const double a1 = 2.0 * alpha / sqrt(M_PI);
    const double a2 = -M_PI * M_PI / (alpha * alpha * L * L);
    const double a3 = 2.0 * M_PI / L;
    const double a4 = 2.0 / (L * L);
    const double a5 = 1.0 / (M_PI * L);
    for (int p = -NN; p <= NN; p++)
        for (int q = -NN; q <= NN; q++)
            for (int s = -NN; s <= NN; s++) {
                const double xx0 = x[0] - L * p, xx1 = x[1] - L * q, xx2 = x[2] - L * s;
                const double ax2 = xx0 * xx0 + xx1 * xx1 + xx2 * xx2;
                const double axx = sqrt(ax2);
                if (axx > 4.0 * L)
                    continue;
                const double t1 = erfc(alpha * axx);
                const double t2 = a1 * exp(-alpha * alpha * ax2);
                const double t = (t1 / axx + t2) / ax2;
                f[0] += t * xx0;
                f[1] += t * xx1;
                f[2] += t * xx2;
                phi -= t1 / axx;
            }
    for (int p = -NN; p <= NN; p++)
        for (int q = -NN; q <= NN; q++)
            for (int s = -NN; s <= NN; s++) {
                const double hh = (double)(p * p + q * q + s * s);
                if (hh > 16 || hh == 0)
                    continue;
                const double hx = p * x[0] + q * x[1] + s * x[2];
                const double e = exp(a2 * hh);
                const double t = a4 * e * sin(a3 * hx) / hh;
                f[0] += t * p;
                f[1] += t * q;
                f[2] += t * s;
                phi -= a5 * e * cos(a3 * hx) / hh;
            }
    const double r = sqrt(r2x), r3 = r * r2x;
    return make_double4(f[0] - x[0] / r3, f[1] - x[1] / r3, f[2] - x[2] / r3, phi + 1.0 / r);
}

// Table node (i, j, k) at r = (i, j, k) h, h = (L/2)/(EWALD_N-1): the correction
// and its partial derivatives by FP64 central differences of step h/64 (T is
// smooth there; the removed Newtonian singularity cancels inside C).
__global__ void k_ewald_table(double L, double4 *__restrict__ t64, float4 *__restrict__ t32) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= EWALD_N * EWALD_N * EWALD_N)
        return;
    const int i = id / (EWALD_N * EWALD_N), j = (id / EWALD_N) % EWALD_N, k = id % EWALD_N;
    const double h = 0.5 * L / (EWALD_N - 1);
    const double x = i * h, y = j * h, z = k * h, d = h / 64;
    double4 out[4];
    out[0] = ewald_point(x, y, z, L);
    for (int b = 0; b < 3; b++) {
        const double4 cp = ewald_point(x + (b == 0) * d, y + (b == 1) * d, z + (b == 2) * d, L);
        const double4 cm = ewald_point(x - (b == 0) * d, y - (b == 1) * d, z - (b == 2) * d, L);
        const double s = 1.0 / (2.0 * d);
        out[1 + b] = make_double4((cp.x - cm.x) * s, (cp.y - cm.y) * s, (cp.z - cm.z) * s, (cp.w - cm.w) * s);
    }
    for (int c = 0; c < 4; c++) {
        t64[4 * (size_t)id + c] = out[c];
        t32[4 * (size_t)id + c] = make_float4((float)out[c].x, (float)out[c].y, (float)out[c].z, (float)out[c].w);
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_ewald_table(double L, double4 *t64, float4 *t32, hipStream_t st) {
    const int n = EWALD_N * EWALD_N * EWALD_N;
    hipLaunchKernelGGL(k_ewald_table, dim3((n + 127) / 128), dim3(128), 0, st, L, t64, t32);
    return hipGetLastError();
}

hipError_t launch_reset_counter(int *c, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU>
static void launch_one(const BatchArgs &A, int persistent_blocks, hipStream_t st) {
    // every XCD needs workgroups to drain its slice: round up to a multiple of NXCD
    persistent_blocks = (persistent_blocks + NXCD - 1) / NXCD * NXCD;
    if (persistent_blocks > 0)
        hipLaunchKernelGGL((k_batch_persistent<T, REL, COMP, SOFT, PER, BU>), dim3(persistent_blocks),
                           dim3(64), 0, st, A, A.bpos, A.cpos, (const T *)A.mom, A.gfirst,
                           A.gcount, A.gcen, A.lpt, A.xoff, A.off, A.idx, (V4<T> *)A.out,
                           A.counter, A.ewald);
    else
        hipLaunchKernelGGL((k_batch<T, REL, COMP, SOFT, PER, BU>), dim3(A.ngroups), dim3(64), 0, st, A);
}

// the unroll factor: a choice for the group-relative compensated FP32 kernels only
template <typename T, bool REL, bool COMP, int SOFT, int PER>
static void launch_bu(const BatchArgs &A, int pb, hipStream_t st) {
    if constexpr (REL && COMP && sizeof(T) == 4) {
        if (A.unroll == 2)
            return launch_one<T, REL, COMP, SOFT, PER, 2>(A, pb, st);
        if (A.unroll == 8)
            return launch_one<T, REL, COMP, SOFT, PER, 8>(A, pb, st);
    }
    launch_one<T, REL, COMP, SOFT, PER, 4>(A, pb, st);
}

template <typename T, bool REL, bool COMP, int SOFT>
static void launch_per(const BatchArgs &A, int per, int pb, hipStream_t st) {
    if (per == 1)
        launch_bu<T, REL, COMP, SOFT, 1>(A, pb, st);
    else if (per == 2)
        launch_bu<T, REL, COMP, SOFT, 2>(A, pb, st);
    else
        launch_bu<T, REL, COMP, SOFT, 0>(A, pb, st);
}

template <typename T, bool REL, bool COMP>
static void launch_soft(const BatchArgs &A, int per, int pb, hipStream_t st) {
    switch (A.soft) {
    case SOFT_K1: launch_per<T, REL, COMP, SOFT_K1>(A, per, pb, st); break;
    case SOFT_PLUMMER: launch_per<T, REL, COMP, SOFT_PLUMMER>(A, per, pb, st); break;
    default: launch_per<T, REL, COMP, SOFT_NONE>(A, per, pb, st);
    }
}

hipError_t launch_batch(const BatchView &v, int variant, int persistent_blocks, hipStream_t st) {
    BatchArgs A;
    A.bpos = v.bpos;
    A.cpos = v.cpos;
    A.mom = variant == VAR_F64 ? (const void *)v.mom64 : (const void *)v.mom32;
    A.gfirst = v.gfirst;
    A.gcount = v.gcount;
    A.gslot_off = v.gslot_off;
    A.gslot = v.gslot;
    A.gcen = v.gcen;
    A.lpt = v.lpt;
    A.xoff = v.xoff;
    A.order_np = v.order_np;
    A.off = v.off;
    A.idx = v.idx;
    A.out = variant == VAR_F64 ? (void *)v.out64 : (void *)v.out32;
    A.counter = v.counter;
    A.ewald = v.ewald32;
    A.ewald64 = v.ewald64;
    A.ngroups = v.ngroups;
    A.soft = v.soft;
    A.eps = v.eps;
    A.G = v.G;
    A.box = v.box;
    A.img = v.img;
    A.img_f = 2 * v.nimage + 1;
    A.unroll = v.unroll;
    A.ewgroup = v.ewgroup;
    if (A.ngroups <= 0)
        return hipSuccess;
    if (v.periodic < 0 || v.periodic > 2 || (v.periodic == 2 && !v.img))
        return hipErrorInvalidValue;
    const int per = v.periodic;
    switch (variant) {
    case VAR_REL32C: launch_soft<float, true, true>(A, per, persistent_blocks, st); break;
    case VAR_REL32: launch_soft<float, true, false>(A, per, persistent_blocks, st); break;
    case VAR_ABS32: launch_soft<float, false, false>(A, per, persistent_blocks, st); break;
    case VAR_F64: launch_soft<double, true, false>(A, per, persistent_blocks, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_group64(const BatchView &v, int g, hipStream_t st) {
    BatchArgs A;
    A.bpos = v.bpos;
    A.cpos = v.cpos;
    A.mom = v.mom64;
    A.gfirst = v.gfirst;
    A.gcount = v.gcount;
    A.gslot_off = v.gslot_off;
    A.gslot = v.gslot;
    A.gcen = v.gcen;
    A.lpt = v.lpt;
    A.xoff = v.xoff;
    A.order_np = v.order_np;
    A.off = v.off;
    A.idx = v.idx;
    A.out = v.out64;
    A.counter = v.counter;
    A.ewald = v.ewald32;
    A.ewald64 = v.ewald64;
    A.ngroups = v.ngroups;
    A.soft = v.soft;
    A.eps = v.eps;
    A.G = v.G;
    A.box = v.box;
    A.img = v.img;
    A.img_f = 2 * v.nimage + 1;
    A.unroll = v.unroll;
    A.ewgroup = v.ewgroup;
    if (v.periodic < 0 || v.periodic > 2 || (v.periodic == 2 && !v.img))
        return hipErrorInvalidValue;
    const int per = v.periodic;
#define G64(S) \
    do { if (per == 1) hipLaunchKernelGGL((k_group64<S, 1>), dim3(1), dim3(64), 0, st, A, g); \
         else if (per == 2) hipLaunchKernelGGL((k_group64<S, 2>), dim3(1), dim3(64), 0, st, A, g); \
         else hipLaunchKernelGGL((k_group64<S, 0>), dim3(1), dim3(64), 0, st, A, g); } while (0)
    switch (v.soft) {
    case SOFT_K1: G64(SOFT_K1); break;
    case SOFT_PLUMMER: G64(SOFT_PLUMMER); break;
    default: G64(SOFT_NONE);
    }
#undef G64
    return hipGetLastError();
}

template <bool PER>
static void launch_direct_soft(dim3 grid, const double4 *bpos, int64_t n, const int64_t *sample,
                               int ns, int soft, double eps, double4 *part, int64_t chunk,
                               const double4 *tab, double box, hipStream_t st) {
    if (soft == SOFT_K1)
        hipLaunchKernelGGL((k_direct64<SOFT_K1, PER>), grid, dim3(64), 0, st, bpos, n, sample, ns,
                           eps, part, chunk, tab, box);
    else if (soft == SOFT_PLUMMER)
        hipLaunchKernelGGL((k_direct64<SOFT_PLUMMER, PER>), grid, dim3(64), 0, st, bpos, n, sample,
                           ns, eps, part, chunk, tab, box);
    else
        hipLaunchKernelGGL((k_direct64<SOFT_NONE, PER>), grid, dim3(64), 0, st, bpos, n, sample,
                           ns, eps, part, chunk, tab, box);
}

hipError_t launch_direct64(const double4 *bpos, int64_t n, const int64_t *sample, int ns, int soft,
                           double eps, double G, double4 *part, int nsplit, double4 *out,
                           const double4 *ewald64, double box, hipStream_t st) {
    if (ns <= 0)
        return hipSuccess;
    const int64_t chunk = (n + nsplit - 1) / nsplit;
    dim3 grid((ns + 63) / 64, nsplit);
    if (ewald64)
        launch_direct_soft<true>(grid, bpos, n, sample, ns, soft, eps, part, chunk, ewald64, box, st);
    else
        launch_direct_soft<false>(grid, bpos, n, sample, ns, soft, eps, part, chunk, nullptr, box, st);
    hipLaunchKernelGGL(k_direct_reduce, dim3((ns + 255) / 256), dim3(256), 0, st, part, ns, nsplit, G, out);
    return hipGetLastError();
}

}  // namespace pgrav
