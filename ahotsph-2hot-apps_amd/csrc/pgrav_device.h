// Device-side building blocks shared by every gravity kernel of libpgrav.
//
// One interaction = one (sink, source) pair.  Sources are order-0 (monopole),
// order-2 (dipole + quadrupole) or order-4 (… + octupole + hexadecapole)
// multipole cells, or bodies.  The arithmetic is the Cartesian traceless
// expansion of nln/pgrav.cu:66-203 (pH) rewritten so that every constant and
// every power of the cell radius R is folded into a "prepared record" once per
// source, leaving ~113 VALU operations per order-4 interaction instead of the
// ~160 of the recursive t *= (2k-1) R / r^2 form.  Same expansion, same
// moments; results agree with the reference formula to FP32 rounding.
//
// Prepared order-4 record (PREC4 = 48 floats, 192 B, read with scalar loads):
//   [0..2]   G1 = R (qx,qy,qz)                       dipole, c1 = 1
//   [3]      m
//   [4..9]   1.5 R^2 (qxx,qxy,qxz,qyy,qyz,qzz)       quadrupole rows, c2 = 1/2
//   [10..24] 5 R^3 x 15 octupole coefficients on the basis
//            A = x^2-z^2, B = y^2-z^2, xy, xz, yz     (rows x,y,z), c3 = 1/3
//   [25..45] 26.25 R^4 x 21 hexadecapole coefficients on the basis
//            U1 = x(x^2-3z^2), U2 = y(y^2-3z^2), U3 = y(x^2-z^2), U4 = z(3x^2-z^2),
//            U5 = x(y^2-z^2), U6 = xyz, U7 = z(3y^2-z^2)   (rows x,y,z), c4 = 1/4
//   [46]     sum m |x - cofm|^2 (the periodic kernels' potential term)
//   [47]     pad
// With Gk the order-k gradient vector and Pk = x.Gk (k = 1..4):
//   S_phi = m + ri2 (P1 + ri2 (P2 + ri2 (P3 + ri2 P4)))
//   T'    = P1 + 2ri2 (P2 + 1.5ri2 (P3 + 4/3 ri2 P4))
//   V     = G1 + 2ri2 (G2 + 1.5ri2 (G3 + 4/3 ri2 G4))
//   phi  += rinv S_phi,   a += rinv ri2 (x (S_phi + 2 ri2 T') - V),   rinv = -1/r
// which is term-for-term nln/pgrav.cu:73-203 (the (2k+1) c_k = 2 + c_k identity
// turns the acceleration's per-order factors into S_phi + 2T).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pgrav_record.h"

namespace pgrav {

// Softening kinds, nln/mac.c:196-237 (only 0 and 4 are coherent in the reference).
enum Soft { SOFT_NONE = 0, SOFT_PLUMMER = 1, SOFT_K1 = 4 };

template <typename T>
struct Acc4T {
    T x, y, z, w;  // ax, ay, az, phi
};
typedef Acc4T<float> Acc4;

// -1/sqrt(r2): v_rsq_f32 in FP32; the FP64 path uses a correctly rounded
// square root and divide (there is no full-rate FP64 rsqrt).
__device__ __forceinline__ float nrsq(float r2) { return -__builtin_amdgcn_rsqf(r2); }
__device__ __forceinline__ double nrsq(double r2) { return -1.0 / sqrt(r2); }

// ---------------------------------------------------------------- interactions
// Order-4 (and order-2 when ORDER == 2) multipole interaction.  M points at a
// prepared record; the caller passes m separately (it may come from LDS).
// QSIGN = -1 reproduces nln/pgrav.cu's "phi -= eqe" for the quadrupole term.
// TM is the record's storage type (T itself, or float when T is a dual number).
template <int ORDER, bool DIPOLE, int QSIGN, typename T, typename TM = T>
__device__ __forceinline__ void cell_interact(const TM *__restrict__ M, T m, T x, T y, T z, T eps2,
                                              Acc4T<T> &a) {
    const T x2 = x * x, y2 = y * y, z2 = z * z;
    const T r2 = x2 + y2 + z2 + eps2;
    const T rinv = nrsq(r2);
    const T ri2 = rinv * rinv;

    // order 1
    T g1x = 0, g1y = 0, g1z = 0, p1 = 0;
    if (DIPOLE) {
        g1x = M[0]; g1y = M[1]; g1z = M[2];
        p1 = g1x * x + g1y * y + g1z * z;
    }
    // order 2
    const T g2x = M[4] * x + M[5] * y + M[6] * z;
    const T g2y = M[5] * x + M[7] * y + M[8] * z;
    const T g2z = M[6] * x + M[8] * y + M[9] * z;
    const T p2 = g2x * x + g2y * y + g2z * z;
    const T rc = T(2) * ri2;
    if (ORDER == 2) {
        T sphi = (QSIGN > 0 ? p2 : -p2);
        sphi = m + ri2 * (p1 + ri2 * sphi);
        const T tp = p1 + rc * p2;
        T sa = sphi + rc * tp;
        if (QSIGN < 0) sa += T(2) * ri2 * ri2 * p2;
        const T ra = rinv * ri2;
        a.w += rinv * sphi;
        if (DIPOLE) {
            const T vx = g1x + rc * g2x, vy = g1y + rc * g2y, vz = g1z + rc * g2z;
            a.x += ra * (x * sa - vx);
            a.y += ra * (y * sa - vy);
            a.z += ra * (z * sa - vz);
        } else {
            // V = rc G2: a += (ra sa) x - (ra rc) G2, two FMAs a component
            const T ras = ra * sa, nrarc = -(ra * rc);
            a.x = fma(nrarc, g2x, fma(ras, x, a.x));
            a.y = fma(nrarc, g2y, fma(ras, y, a.y));
            a.z = fma(nrarc, g2z, fma(ras, z, a.z));
        }
        return;
    }
    // order 3: basis A, B, xy, xz, yz
    const T A = x2 - z2, B = y2 - z2;
    const T xy = x * y, xz = x * z, yz = y * z;
    const TM *o = M + 10;
    const T g3x = o[0] * A + o[1] * B + o[2] * xy + o[3] * xz + o[4] * yz;
    const T g3y = o[5] * A + o[6] * B + o[7] * xy + o[8] * xz + o[9] * yz;
    const T g3z = o[10] * A + o[11] * B + o[12] * xy + o[13] * xz + o[14] * yz;
    const T p3 = g3x * x + g3y * y + g3z * z;
    // order 4: basis U1..U7
    const T z3 = T(3) * z2;
    const T U1 = x * (x2 - z3);
    const T U2 = y * (y2 - z3);
    const T U3 = y * A;
    const T U4 = z * (T(3) * x2 - z2);
    const T U5 = x * B;
    const T U6 = xy * z;
    const T U7 = z * (T(3) * y2 - z2);
    const TM *h = M + 25;
    const T g4x = h[0] * U1 + h[1] * U2 + h[2] * U3 + h[3] * U4 + h[4] * U5 + h[5] * U6 + h[6] * U7;
    const T g4y = h[7] * U1 + h[8] * U2 + h[9] * U3 + h[10] * U4 + h[11] * U5 + h[12] * U6 + h[13] * U7;
    const T g4z = h[14] * U1 + h[15] * U2 + h[16] * U3 + h[17] * U4 + h[18] * U5 + h[19] * U6 + h[20] * U7;
    const T p4 = g4x * x + g4y * y + g4z * z;

    const T rb = T(1.5) * ri2, rq = T(4.0 / 3.0) * ri2;
    T sphi = p3 + ri2 * p4;
    sphi = (QSIGN > 0 ? p2 : -p2) + ri2 * sphi;
    sphi = m + ri2 * (p1 + ri2 * sphi);
    const T tp = p1 + rc * (p2 + rb * (p3 + rq * p4));
    T sa = sphi + rc * tp;
    if (QSIGN < 0) sa += T(2) * ri2 * ri2 * p2;
    const T ra = rinv * ri2;
    a.w += rinv * sphi;
    // W = G2 + rb (G3 + rq G4), V = G1 + rc W
    const T wx = g2x + rb * (g3x + rq * g4x);
    const T wy = g2y + rb * (g3y + rq * g4y);
    const T wz = g2z + rb * (g3z + rq * g4z);
    if (DIPOLE) {
        a.x += ra * (x * sa - (g1x + rc * wx));
        a.y += ra * (y * sa - (g1y + rc * wy));
        a.z += ra * (z * sa - (g1z + rc * wz));
    } else {
        // a += (ra sa) x - (ra rc) W: two FMAs a component, no zero dipole added
        const T ras = ra * sa, nrarc = -(ra * rc);
        a.x = fma(nrarc, wx, fma(ras, x, a.x));
        a.y = fma(nrarc, wy, fma(ras, y, a.y));
        a.z = fma(nrarc, wz, fma(ras, z, a.z));
    }
}

// Monopole / body with Plummer softening eps2 (0 = none), nln/pgrav.cu:380-400,
// nln/do_grav_sse4.c:2188-2215.
template <typename T>
__device__ __forceinline__ void mono_interact(T m, T x, T y, T z, T eps2, Acc4T<T> &a) {
    // eps2 seeds the sum: three FMAs
    const T r2 = fma(x, x, fma(y, y, fma(z, z, eps2)));
    const T rinv = nrsq(r2);
    const T mr = m * rinv;
    const T mr3 = mr * rinv * rinv;
    a.w += mr;
    a.x += x * mr3;
    a.y += y * mr3;
    a.z += z * mr3;
}

// Body pair with Dehnen K1 softening (nln/pgrav.cu:745-770).  eps_inv = 1/eps,
// eps2 = eps^2.
template <typename T>
__device__ __forceinline__ void k1_interact(T m, T x, T y, T z, T eps_inv, T eps2, Acc4T<T> &a) {
    const T r2 = x * x + y * y + z * z;
    T phi, eqe;
    if (r2 > eps2) {
        const T rinv = nrsq(r2);
        phi = m * rinv;
        eqe = phi * rinv * rinv;
    } else {
        const T eps_inv2 = eps_inv * eps_inv;
        const T u2 = r2 * eps_inv2 - T(1);
        T t = (((T(45.0 / 32.0)) * u2 + T(-3.0 / 8.0)) * u2 + T(0.5)) * u2 - T(1);
        phi = t * m * eps_inv;
        eqe = ((T(-135.0 / 16.0) * u2 + T(1.5)) * u2 - T(1)) * (m * eps_inv * eps_inv2);
    }
    a.w += phi;
    a.x += x * eqe;
    a.y += y * eqe;
    a.z += z * eqe;
}

// ---------------------------------------------------------------- wave sums
// The sum of a value over the 64 lanes of a wave by DPP row operations (VALU
// only, no LDS): pairs and quads by quad permutes, 8 and 16 lanes by the half-
// row and row mirrors, then rows 0+1 and 2+3 by the row-15 broadcast and all
// four by the row-31 broadcast, which leaves the total in lane 63, read back
// into every lane.  The lane order of the additions is fixed, so the result is
// reproducible.  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __builtin_bit_cast(float, dpp_mov<CTRL, ROWS>(__builtin_bit_cast(int, v)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = dpp_mov<CTRL, ROWS>((int)(b & 0xffffffffll));
    const int hi = dpp_mov<CTRL, ROWS>((int)(b >> 32));
    return v + __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
    v = dpp_add<0x140, 0xf>(v);  // row_mirror
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (unsigned int)lo);
    }
}

// Error-free transformation accumulate (Knuth 2Sum): (s, c) += v.
template <typename T>
__device__ __forceinline__ void two_sum_add(T &s, T &c, T v) {
    const T t = s + v;
    const T bp = t - s;
    const T err = (s - (t - bp)) + (v - bp);
    s = t;
    c += err;
}

}  // namespace pgrav
