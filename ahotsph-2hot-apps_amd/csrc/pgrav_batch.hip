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
#include "pgrav_ewald.h"

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
    const float *ewald;  // Ewald correction table (periodic = 1, pgrav_ewald.h)
    const double *ewald64;  // its FP64 original, read by the FP64 kernels
    const uint8_t *img;  // explicit images: the image of each entry, parallel to idx (or null)
    int img_f;           // 2 nimage + 1 images per side
    int unroll;          // order-0 / body loop unroll of the rel32c kernels: 2, 4 or 8
    int tile;            // list entries staged in LDS per round (rel32c, not periodic 1): 32, 64, 128
    double ewratio;      // periodic 1: expansion about the group centre when radius <= ewratio x distance
    int ngroups;
    int soft;
    double eps;
    double G;
    double box;
    double ibox;  // 1 / box
};

// nearest periodic image of a displacement (identity when not periodic)
template <bool PERIODIC, typename T>
__device__ __forceinline__ T pbc(T x, T box, T ibox) {
    if (PERIODIC)
        return x - box * rint(x * ibox);
    return x;
}

// the table a kernel reads: FP32 arithmetic its FP32 copy, FP64 the FP64 original
__device__ __forceinline__ const float *ewald_table(const BatchArgs &A, float) {
    return A.ewald;
}
__device__ __forceinline__ const double *ewald_table(const BatchArgs &A, double) {
    return A.ewald64;
}

// Periodic (nearest image + Ewald correction) per group, in LDS: each lane's
// running sums of the expansion terms (20, below) of the entries it loaded, and
// the group's sums.  (Which entries of a tile take the correction per pair is a
// wave-uniform 64-bit mask in scalar registers.)
template <typename T, int PER>
struct EwShared {
    static constexpr int NP = PER == 1 ? 20 : 1;
    T part[NP][64];
    double acc[NP];
};

// The correction summed over a group's expanded sources (group radius small
// against the distance to their nearest other images, no sink seeing another
// image than the centre does), as a field of the sink's offset s from the
// group centre: with r0 each source's offset from the centre and m its mass,
//   P = sum m psi(r0), A = sum m C(r0), B = sum m J(r0), Tm = sum m T(r0)
// (a cell adds its traceless second moment's coupling, (1/2) J:Q to P and
// (1/2) T:Q to A), the sink at s = r0 - r takes
//   C(s) = A - B s + Tm s s / 2,  psi(s) = P - A.s + s.B.s / 2 - Tm s s s / 6.
// acc: [0] P, [1..3] A, [4..9] B (xx xy xz yy yz zz), [10..19] Tm.
template <typename T>
__device__ __forceinline__ void ewald_expansion_apply(const double *acc, T sx, T sy, T sz,
                                                      Acc4T<T> &a) {
    T t[10];
#pragma unroll
    for (int q = 0; q < 10; q++) t[q] = T(acc[10 + q]);
    const T bxx = T(acc[4]), bxy = T(acc[5]), bxz = T(acc[6]), byy = T(acc[7]), byz = T(acc[8]),
            bzz = T(acc[9]);
    const T ax = T(acc[1]), ay = T(acc[2]), az = T(acc[3]);
    const T txx = t[0] * sx + t[1] * sy + t[2] * sz, txy = t[1] * sx + t[3] * sy + t[4] * sz,
            txz = t[2] * sx + t[4] * sy + t[5] * sz, tyy = t[3] * sx + t[6] * sy + t[7] * sz,
            tyz = t[4] * sx + t[7] * sy + t[8] * sz, tzz = t[5] * sx + t[8] * sy + t[9] * sz;
    const T ux = txx * sx + txy * sy + txz * sz, uy = txy * sx + tyy * sy + tyz * sz,
            uz = txz * sx + tyz * sy + tzz * sz;  // Tm s s
    const T vx = bxx * sx + bxy * sy + bxz * sz, vy = bxy * sx + byy * sy + byz * sz,
            vz = bxz * sx + byz * sy + bzz * sz;  // B s
    a.x += ax - vx + T(0.5) * ux;
    a.y += ay - vy + T(0.5) * uy;
    a.z += az - vz + T(0.5) * uz;
    a.w += T(acc[0]) - (ax * sx + ay * sy + az * sz) + T(0.5) * (vx * sx + vy * sy + vz * sz) -
           T(1.0 / 6.0) * (ux * sx + uy * sy + uz * sz);
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
__device__ __forceinline__ V4<T> source_offset(double4 c, double4 gc, double box, double ibox) {
    double dx = c.x, dy = c.y, dz = c.z;
    if (REL || PERIODIC) {
        dx -= gc.x;
        dy -= gc.y;
        dz -= gc.z;
    }
    if (PERIODIC) {
        // (a product with 1 / box, not a division: the image of a source exactly half
        // a box away is a tie either way)
        dx -= box * rint(dx * ibox);
        dy -= box * rint(dy * ibox);
        dz -= box * rint(dz * ibox);
        if (!REL) {  // absolute variant: re-add the group centre after wrapping
            dx += gc.x;
            dy += gc.y;
            dz += gc.z;
        }
    }
    return V4<T>{T(dx), T(dy), T(dz), T(c.w)};
}

// The Ewald correction of one tile entry (periodic = 1), in the tile loader:
// either the entry joins the group's expansion -- its (psi, C, J, T) at r0 =
// source - centre, times its mass, plus a cell's second-moment coupling -- or
// it is marked for per-pair lookups (the return value).  `on` says whether this
// lane holds an entry; the terms go to the lane's own running sums in LDS.
template <typename T, int PER>
__device__ __forceinline__ bool ewald_tile_entry(const BatchArgs &A, EwShared<T, PER> *ew, bool on,
                                                 int lane, double4 src, double4 gc, bool cell,
                                                 const T *rec) {
    T v[20];
    bool pair = false;
    if (on) {
        double dx = src.x - gc.x, dy = src.y - gc.y, dz = src.z - gc.z;
        dx -= A.box * rint(dx * A.ibox);
        dy -= A.box * rint(dy * A.ibox);
        dz -= A.box * rint(dz * A.ibox);
        const double am = fmax(fabs(dx), fmax(fabs(dy), fabs(dz)));
        // the expansion holds when every sink sees the centre's image of the source
        // and the source's other images are far against the group radius
        const bool expand = am < 0.5 * A.box - gc.w && gc.w <= A.ewratio * (A.box - am);
        pair = !expand;
#pragma unroll
        for (int q = 0; q < 20; q++) v[q] = T(0);
        if (expand) {
            EwField<T> F;
            ewald_field<T>(ewald_table(A, T(0)), T(dx), T(dy), T(dz), T(A.box), F);
            const T m = T(src.w);
            v[0] = m * F.psi;
#pragma unroll
            for (int q = 0; q < 3; q++) v[1 + q] = m * F.c[q];
#pragma unroll
            for (int q = 0; q < 6; q++) v[4 + q] = m * F.j[q];
#pragma unroll
            for (int q = 0; q < 10; q++) v[10 + q] = m * F.t[q];
            if (cell) {
                // traceless second moment Q = sum m (d d - |d|^2/3) from the prepared
                // record (1.5 Q, order xx xy xz yy yz zz)
                const T k = T(2.0 / 3.0);
                const T qxx = k * rec[4], qxy = k * rec[5], qxz = k * rec[6], qyy = k * rec[7],
                        qyz = k * rec[8], qzz = k * rec[9];
                const T h = T(0.5), two = T(2);
                const T *j = F.j, *t = F.t;
                v[0] += h * (j[0] * qxx + j[3] * qyy + j[5] * qzz +
                             two * (j[1] * qxy + j[2] * qxz + j[4] * qyz));
                v[1] += h * (t[0] * qxx + t[3] * qyy + t[5] * qzz + two * (t[1] * qxy + t[2] * qxz + t[4] * qyz));
                v[2] += h * (t[1] * qxx + t[6] * qyy + t[8] * qzz + two * (t[3] * qxy + t[4] * qxz + t[7] * qyz));
                v[3] += h * (t[2] * qxx + t[7] * qyy + t[9] * qzz + two * (t[4] * qxy + t[5] * qxz + t[8] * qyz));
            }
        }
        // this lane's running sums (summed over the lanes once per group)
#pragma unroll
        for (int q = 0; q < 20; q++) ew->part[q][lane] += v[q];
    }
    return pair;
}

// PER: 0 isolated, 1 nearest image + Ewald correction, 2 explicit images
// BU: unroll factor of the order-0 and body loops (2, 4 or 8; the autotuner picks it
// for the group-relative compensated FP32 kernels, the others use 4)
// TILE: list entries staged in LDS per round (32, 64 or 128 -- one barrier pair and
// one compensated fold per round; 64 in the periodic-1 kernels, whose per-entry
// Ewald state is a lane's)
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU, int TILE = 64>
__device__ __forceinline__ void eval_group(const BatchArgs &A, int g, V4<T> *lds,
                                           EwShared<T, PER> *ew) {
    static_assert(TILE == 32 || TILE == 64 || TILE == 128, "tile of 32, 64 or 128 entries");
    static_assert(PER != 1 || TILE == 64, "the periodic-1 loader keeps one entry per lane");
    constexpr bool PERIODIC = PER == 1;
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
    if (PERIODIC) {
#pragma unroll
        for (int q = 0; q < 20; q++) ew->part[q][lane] = T(0);
    }

#pragma unroll 1
    for (int kind = 0; kind < 4; kind++) {
        const int64_t *offk = A.off + (int64_t)kind * (A.ngroups + 1);
        const uint32_t *__restrict__ idx = A.idx;
        const int64_t o0 = offk[g], o1 = offk[g + 1];
        for (int64_t t0 = o0; t0 < o1; t0 += TILE) {
            const int nt = (int)min((int64_t)TILE, o1 - t0);
            uint64_t pmask = 0;  // periodic 1: bit k set when entry k takes the per-pair correction
#pragma unroll
            for (int h = 0; h < (TILE + 63) / 64; h++) {
                const int u = lane + 64 * h;  // the tile entry this lane loads
                double4 src = make_double4(0, 0, 0, 0);
                uint32_t e = 0;
                if (u < nt) {
                    e = idx[t0 + u];
                    src = kind == 3 ? A.bpos[e] : A.cpos[e];
                    if (PER == 2) {
                        // image (i, j, k) in nln/mac.c:91-111 order: the source sits at
                        // x + (i, j, k) box, added in FP64
                        const int im = A.img[t0 + u], f = A.img_f, hf = f >> 1;
                        src.x += (double)(im / (f * f) - hf) * A.box;
                        src.y += (double)((im / f) % f - hf) * A.box;
                        src.z += (double)(im % f - hf) * A.box;
                    }
                    lds[u] = source_offset<T, REL, PERIODIC>(src, gc, A.box, A.ibox);
                    if (PERIODIC && kind < 3)
                        trace += (double)mom[(size_t)e * PREC4 + PREC_TRACE];
                }
                if (PERIODIC) {
                    const bool pair = ewald_tile_entry<T, PER>(A, ew, u < nt, lane, src, gc,
                                                               kind < 3, mom + (size_t)e * PREC4);
                    // one entry per lane (TILE 64): the wave's flags as a scalar mask
                    pmask = __builtin_amdgcn_ballot_w64(pair);
                }
            }
            __syncthreads();
            Acc4T<T> p = {0, 0, 0, 0};
            // Offsets from the sink: an expanded source's image is the centre's for
            // every sink of the group (its offset from the centre was wrapped in the
            // loader and no sink crosses a half box from it), so only the sources
            // taken per pair need the nearest image per sink.
#define PGRAV_OFFSET(r, x, y, z, pp)                                \
    const bool pp = PERIODIC && ((pmask >> k) & 1);                 \
    T x = sx - r.x, y = sy - r.y, z = sz - r.z;                     \
    if (pp) {                                                       \
        x = pbc<PERIODIC>(x, box, ibox);                            \
        y = pbc<PERIODIC>(y, box, ibox);                            \
        z = pbc<PERIODIC>(z, box, ibox);                            \
    }
            if (kind == 0) {
                for (int k = 0; k < nt; k++) {
                    const uint32_t ek = idx[t0 + k];
                    const V4<T> r = lds[k];
                    PGRAV_OFFSET(r, x, y, z, pp)
                    cell_interact<4, false, 1, T>(mom + (size_t)ek * PREC4, r.w, x, y, z, eps2_cell, p);
                    if (pp)
                        ewald_pair<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else if (kind == 1) {
                for (int k = 0; k < nt; k++) {
                    const uint32_t ek = idx[t0 + k];
                    const V4<T> r = lds[k];
                    PGRAV_OFFSET(r, x, y, z, pp)
                    cell_interact<2, false, 1, T>(mom + (size_t)ek * PREC4, r.w, x, y, z, eps2_cell, p);
                    if (pp)
                        ewald_pair<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else if (kind == 2) {
#pragma unroll BU
                for (int k = 0; k < nt; k++) {
                    const V4<T> r = lds[k];
                    PGRAV_OFFSET(r, x, y, z, pp)
                    mono_interact<T>(r.w, x, y, z, eps2_cell, p);
                    if (pp)
                        ewald_pair<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            } else {
                // bodies: ~18 VALU per pair, so unroll to keep several LDS reads in flight
#pragma unroll BU
                for (int k = 0; k < nt; k++) {
                    const V4<T> r = lds[k];
                    PGRAV_OFFSET(r, x, y, z, pp)
                    if (SOFT == SOFT_K1) {
                        k1_interact<T>(r.w, x, y, z, eps_inv, eps2, p);
                    } else if (SOFT == SOFT_PLUMMER) {
                        mono_interact<T>(r.w, x, y, z, eps2, p);
                    } else {  // unsoftened: drop the self pair
                        const bool self = x == T(0) && y == T(0) && z == T(0);
                        mono_interact<T>(self ? T(0) : r.w, x, y, z, self ? T(1) : T(0), p);
                    }
                    // the self pair too: at r = 0 the correction is its limit, C = 0 and
                    // psi(0) = 2.8373/L, the body's potential from its own images and the
                    // background
                    if (pp)
                        ewald_pair<T>(ewald_table(A, T(0)), r.w, x, y, z, box, p);
                }
            }
#undef PGRAV_OFFSET
            __syncthreads();
            fold<T>(s, c, p, COMP);
        }
    }
    if (PERIODIC) {
        // the lanes' running sums, summed per term in a fixed order (the last
        // tile's barrier has passed)
        if (lane < 20) {
            double t = 0.0;
            for (int k = 0; k < 64; k++) t += (double)ew->part[lane][k];
            ew->acc[lane] = t;
        }
        __syncthreads();
        Acc4T<T> p = {0, 0, 0, 0};
        // the expanded sources' correction at this sink's offset from the centre
        ewald_expansion_apply<T>(ew->acc, T(sp.x - gc.x), T(sp.y - gc.y), T(sp.z - gc.z), p);
        // the cells' corrections were taken at their centres of mass; the second-
        // order remainder of the potential is (1/2) H : M2 summed over the cells,
        // with H the correction's Hessian (trace -4 pi / L^3 everywhere) and M2 the
        // cell's second moments: its trace part, -(2 pi / 3) / L^3 sum m |x - cofm|^2,
        // is added here, the traceless part with the expansion (or left out for the
        // cells taken per pair)
        const double tr = wave_sum(trace);
        const double b3 = A.box * A.box * A.box;
        p.w += T(-2.0 * M_PI / 3.0 / b3 * tr);
        fold<T>(s, c, p, COMP);
    }
    if (live) {
        const T Gn = T(A.G);  // nln/mac.c:269-277: results scaled by G at the end
        V4<T> r = {Gn * (s.x + c.x), Gn * (s.y + c.y), Gn * (s.z + c.z), Gn * (s.w + c.w)};
        ((V4<T> *)A.out)[slot] = r;
    }
}

// Minimum waves per SIMD asked of the register allocator (FP32 kernels).  The
// periodic kernels' tile loader evaluates the Ewald table (20 values, a Taylor
// step, the symmetry) beside the pair loop's live state.
#ifndef PGRAV_MINWAVES_FP32
#define PGRAV_MINWAVES_FP32 8  // the FP32 kernels without the Ewald table, unroll <= 4: left
                               // alone they take 106 SGPRs (the scalar-loaded records) and 7
                               // waves; held to 8 (78 SGPRs, a few spilled to VGPR lanes, no
                               // scratch) the clustered 16.8 M launch runs 1.5 % faster.  Unroll
                               // 8 would spill VGPRs to scratch at 8 waves and keeps the default.
#endif
#ifndef PGRAV_MINWAVES_PER1
#define PGRAV_MINWAVES_PER1 7  // periodic 1: 72 VGPRs, 3 spilled once per group (38.7 ms at
                               // the 256^3 lattice against 38.9 with 6 waves and 78 VGPRs)
#endif
template <typename T, int PER, int BU>
struct MinWaves {
    static constexpr int value =
        sizeof(T) == 4 ? (PER == 1 ? PGRAV_MINWAVES_PER1 : (BU >= 8 ? 1 : PGRAV_MINWAVES_FP32)) : 1;
};

// one workgroup per group, groups taken in longest-first order
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU, int TILE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MinWaves<T, PER, BU>::value)))
void k_batch(BatchArgs A) {
    __shared__ V4<T> lds[TILE];
    __shared__ EwShared<T, PER> ew;
    eval_group<T, REL, COMP, SOFT, PER, BU, TILE>(A, A.order_np[blockIdx.x], lds, &ew);
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
template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU, int TILE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MinWaves<T, PER, BU>::value)))
void k_batch_persistent(
    BatchArgs A, const double4 *__restrict__ bpos, const double4 *__restrict__ cpos,
    const T *__restrict__ mom, const int64_t *__restrict__ gfirst, const int *__restrict__ gcount,
    const double4 *__restrict__ gcen, const int *__restrict__ lpt, const int *__restrict__ xoff,
    const int64_t *__restrict__ off, const uint32_t *__restrict__ idx, V4<T> *__restrict__ out,
    int *__restrict__ counter, const float *__restrict__ ewald) {
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
    __shared__ V4<T> lds[TILE];
    __shared__ EwShared<T, PER> ew;
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
        eval_group<T, REL, COMP, SOFT, PER, BU, TILE>(A, A.lpt[item], lds, &ew);
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
    __shared__ EwShared<double, PER> ew;
    eval_group<double, false, false, SOFT, PER, 4>(A, g, lds, &ew);
}

__global__ void k_reset_counter(int *c) {
    if (threadIdx.x < 2 + NXCD)
        c[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- FP64 direct sum
// Sampled FP64 direct sum (pattern of nln/grav_n2.c:162-173): grid.y splits the
// sources; partial sums go to part[split][sink] and are reduced in fixed order.
// Periodic: nearest image of every source plus the Ewald correction per pair
// from the FP64 table, so the difference to the tree result is the tree's own
// approximation error (and its expansion of the correction about the groups).
template <int SOFT, bool PERIODIC>
__global__ __launch_bounds__(64) void k_direct64(const double4 *__restrict__ bpos, int64_t n,
                                                 const int64_t *__restrict__ sample, int ns,
                                                 double eps, double4 *__restrict__ part,
                                                 int64_t chunk, const double *__restrict__ tab,
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
                ewald_pair<double>(tab, q.w, x, y, z, box, a);
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
// The correction at one point x in closed form, FP64: the lattice sums of
// nln/ewald.c:13-64 (alpha = 2/L, |n| <= 4 within 4L, |h|^2 <= 16) and their
// derivatives term by term, with the nearest Newtonian term taken out.
// out = psi, C (3), J (xx xy xz yy yz zz), T (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz).
// Real space, y = x - nL, rho = |y|, t1 = erfc(alpha rho), t2 = a1 exp(-alpha^2 rho^2):
//   g = t1/rho + t2, g' = -t2/rho - t1/rho^2 - 2 alpha^2 rho t2,
//   g'' = 2 t2/rho^2 + 2 t1/rho^3 + 4 alpha^4 rho^2 t2, t = g/rho^2,
//   t' = g'/rho^2 - 2 g/rho^3, t'' = g''/rho^2 - 4 g'/rho^3 + 6 g/rho^4,
//   u = t'/rho, u' = t''/rho - t'/rho^2:
//   C_a += t y_a, J_ab += t d_ab + u y_a y_b,
//   T_abc += u (d_ab y_c + d_ac y_b + d_bc y_a) + (u'/rho) y_a y_b y_c, psi -= t1/rho.
// Fourier space, e = exp(a2 |h|^2), s, c = sin, cos(a3 h.x):
//   C += a4 e s/|h|^2 h, J += a3 a4 e c/|h|^2 h h, T -= a3^2 a4 e s/|h|^2 h h h,
//   psi -= a5 e c/|h|^2.
// Newtonian term: C -= x/r^3, J -= d/r^3 - 3 x x/r^5,
//   T += 3 (d_ab x_c + d_ac x_b + d_bc x_a)/r^5 - 15 x x x/r^7, psi += 1/r.
// At x = 0 the n = 0 real-space term minus the Newtonian one is its limit:
// a1 in psi, -(2/3) alpha^2 a1 on J's diagonal, nothing in C and T.
__constant__ int EW_SYM3[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                   {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
__constant__ int EW_SYM2[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};

__device__ void ewald_closed(const double x[3], double L, double out[EW_REC]) {
    const int NN = 4;
    const double alpha = 2.0 / L;
    const double a1 = 2.0 * alpha / sqrt(M_PI);
    const double a2 = -M_PI * M_PI / (alpha * alpha * L * L);
    const double a3 = 2.0 * M_PI / L;
    const double a4 = 2.0 / (L * L);
    const double a5 = 1.0 / (M_PI * L);
    const bool origin = x[0] == 0.0 && x[1] == 0.0 && x[2] == 0.0;
    double psi = M_PI / (alpha * alpha * L * L * L), c[3] = {0, 0, 0}, j[6] = {0, 0, 0, 0, 0, 0};
    double t[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = -NN; p <= NN; p++)
        for (int q = -NN; q <= NN; q++)
            for (int r = -NN; r <= NN; r++) {
                if (origin && p == 0 && q == 0 && r == 0) {
                    psi += a1;
                    j[0] -= 2.0 / 3.0 * alpha * alpha * a1;
                    j[3] -= 2.0 / 3.0 * alpha * alpha * a1;
                    j[5] -= 2.0 / 3.0 * alpha * alpha * a1;
                    continue;
                }
                const double y[3] = {x[0] - L * p, x[1] - L * q, x[2] - L * r};
                const double rho2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
                const double rho = sqrt(rho2);
                if (rho > 4.0 * L)
                    continue;
                const double t1 = erfc(alpha * rho);
                const double t2 = a1 * exp(-alpha * alpha * rho2);
                const double g = t1 / rho + t2;
                const double gp = -t2 / rho - t1 / rho2 - 2.0 * alpha * alpha * rho * t2;
                const double gpp = 2.0 * t2 / rho2 + 2.0 * t1 / (rho2 * rho) +
                                   4.0 * alpha * alpha * alpha * alpha * rho2 * t2;
                const double tt = g / rho2;
                const double tp = gp / rho2 - 2.0 * g / (rho2 * rho);
                const double tpp = gpp / rho2 - 4.0 * gp / (rho2 * rho) + 6.0 * g / (rho2 * rho2);
                const double u = tp / rho, up = (tpp / rho - tp / rho2) / rho;
                for (int a = 0; a < 3; a++) c[a] += tt * y[a];
                for (int k = 0; k < 6; k++) {
                    const int a = EW_SYM2[k][0], b = EW_SYM2[k][1];
                    j[k] += (a == b ? tt : 0.0) + u * y[a] * y[b];
                }
                for (int k = 0; k < 10; k++) {
                    const int a = EW_SYM3[k][0], b = EW_SYM3[k][1], cc = EW_SYM3[k][2];
                    t[k] += u * ((a == b) * y[cc] + (a == cc) * y[b] + (b == cc) * y[a]) +
                            up * y[a] * y[b] * y[cc];
                }
                psi -= t1 / rho;
            }
    for (int p = -NN; p <= NN; p++)
        for (int q = -NN; q <= NN; q++)
            for (int r = -NN; r <= NN; r++) {
                const double hh = (double)(p * p + q * q + r * r);
                if (hh > 16 || hh == 0)
                    continue;
                const double h[3] = {(double)p, (double)q, (double)r};
                const double hx = p * x[0] + q * x[1] + r * x[2];
                const double e = exp(a2 * hh) / hh;
                double sn, cs;
                sincos(a3 * hx, &sn, &cs);
                for (int a = 0; a < 3; a++) c[a] += a4 * e * sn * h[a];
                for (int k = 0; k < 6; k++) j[k] += a3 * a4 * e * cs * h[EW_SYM2[k][0]] * h[EW_SYM2[k][1]];
                for (int k = 0; k < 10; k++)
                    t[k] -= a3 * a3 * a4 * e * sn * h[EW_SYM3[k][0]] * h[EW_SYM3[k][1]] * h[EW_SYM3[k][2]];
                psi -= a5 * e * cs;
            }
    if (!origin) {
        const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        const double rr = sqrt(r2), r3 = rr * r2, r5 = r3 * r2, r7 = r5 * r2;
        for (int a = 0; a < 3; a++) c[a] -= x[a] / r3;
        for (int k = 0; k < 6; k++) {
            const int a = EW_SYM2[k][0], b = EW_SYM2[k][1];
            j[k] += -(a == b ? 1.0 / r3 : 0.0) + 3.0 * x[a] * x[b] / r5;
        }
        for (int k = 0; k < 10; k++) {
            const int a = EW_SYM3[k][0], b = EW_SYM3[k][1], cc = EW_SYM3[k][2];
            t[k] += 3.0 * ((a == b) * x[cc] + (a == cc) * x[b] + (b == cc) * x[a]) / r5 -
                    15.0 * x[a] * x[b] * x[cc] / r7;
        }
        psi += 1.0 / rr;
    }
    out[0] = psi;
    for (int a = 0; a < 3; a++) out[1 + a] = c[a];
    for (int k = 0; k < 6; k++) out[4 + k] = j[k];
    for (int k = 0; k < 10; k++) out[10 + k] = t[k];
}

// one thread per node i >= j >= k of the octant grid
__global__ void k_ewald_table(double L, double *__restrict__ t64, float *__restrict__ t32) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= EWALD_NODES)
        return;
    // invert id = i(i+1)(i+2)/6 + j(j+1)/2 + k
    int i = 0;
    while (ew_node(i + 1, 0, 0) <= id) i++;
    int rem = id - ew_node(i, 0, 0), j = 0;
    while ((j + 1) * (j + 2) / 2 <= rem) j++;
    const int k = rem - j * (j + 1) / 2;
    const double h = 0.5 * L / (EWALD_N - 1);
    const double x[3] = {i * h, j * h, k * h};
    double out[EW_REC];
    ewald_closed(x, L, out);
    for (int q = 0; q < EW_REC; q++) {
        t64[(size_t)id * EW_REC + q] = out[q];
        t32[(size_t)id * EW_REC + q] = (float)out[q];
    }
}

// ---------------------------------------------------------------- launchers
hipError_t launch_ewald_table(double L, double *t64, float *t32, hipStream_t st) {
    hipLaunchKernelGGL(k_ewald_table, dim3((EWALD_NODES + 63) / 64), dim3(64), 0, st, L, t64, t32);
    return hipGetLastError();
}

hipError_t launch_reset_counter(int *c, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

template <typename T, bool REL, bool COMP, int SOFT, int PER, int BU, int TILE = 64>
static void launch_one(const BatchArgs &A, int persistent_blocks, hipStream_t st) {
    // every XCD needs workgroups to drain its slice: round up to a multiple of NXCD
    persistent_blocks = (persistent_blocks + NXCD - 1) / NXCD * NXCD;
    if (persistent_blocks > 0)
        hipLaunchKernelGGL((k_batch_persistent<T, REL, COMP, SOFT, PER, BU, TILE>),
                           dim3(persistent_blocks), dim3(64), 0, st, A, A.bpos, A.cpos,
                           (const T *)A.mom, A.gfirst, A.gcount, A.gcen, A.lpt, A.xoff, A.off,
                           A.idx, (V4<T> *)A.out, A.counter, A.ewald);
    else
        hipLaunchKernelGGL((k_batch<T, REL, COMP, SOFT, PER, BU, TILE>), dim3(A.ngroups), dim3(64),
                           0, st, A);
}

// the unroll factor and the tile: choices for the group-relative compensated FP32
// kernels only (tiles of 32 and 128 entries with the unroll of 4, outside periodic 1)
template <typename T, bool REL, bool COMP, int SOFT, int PER>
static void launch_bu(const BatchArgs &A, int pb, hipStream_t st) {
    if constexpr (REL && COMP && sizeof(T) == 4) {
        if constexpr (PER != 1) {
            if (A.tile == 32)
                return launch_one<T, REL, COMP, SOFT, PER, 4, 32>(A, pb, st);
            if (A.tile == 128)
                return launch_one<T, REL, COMP, SOFT, PER, 4, 128>(A, pb, st);
        }
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
    A.ibox = 1.0 / v.box;
    A.img = v.img;
    A.img_f = 2 * v.nimage + 1;
    A.unroll = v.unroll;
    A.tile = v.tile;
    A.ewratio = v.ewratio;
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
    A.ibox = 1.0 / v.box;
    A.img = v.img;
    A.img_f = 2 * v.nimage + 1;
    A.unroll = v.unroll;
    A.tile = v.tile;
    A.ewratio = v.ewratio;
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
                               const double *tab, double box, hipStream_t st) {
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
                           const double *ewald64, double box, hipStream_t st) {
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
