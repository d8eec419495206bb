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
//   [46..47] pad
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

#define PREC4 48
#define PREC2 12

namespace pgrav {

// Softening kinds, nln/mac.c:196-237 (only 0 and 4 are coherent in the reference).
enum Soft { SOFT_NONE = 0, SOFT_PLUMMER = 1, SOFT_K1 = 4 };

struct Acc4 {
    float x, y, z, w;  // ax, ay, az, phi
};

__device__ __forceinline__ float frsq(float r2) { return __builtin_amdgcn_rsqf(r2); }

// ---------------------------------------------------------------- record prep
// Converts one reference source record (fields in nln/pgrav.cu:16-47 order,
// nfield = 4, 13 or 29) into the prepared layout above.  Runs once per source.
__device__ __host__ inline void prep_record(const float *q, int order, int dipole,
                                            float *out /* PREC4 */) {
    for (int k = 0; k < PREC4; k++) out[k] = 0.0f;
    out[3] = q[0];
    if (order == 0)
        return;
    const float R = q[4];
    if (dipole) {
        out[0] = R * q[5];
        out[1] = R * q[6];
        out[2] = R * q[7];
    }
    const float qxx = q[8], qxy = q[9], qyy = q[10], qxz = q[11], qyz = q[12];
    const float k2 = 1.5f * R * R;
    out[4] = k2 * qxx;
    out[5] = k2 * qxy;
    out[6] = k2 * qxz;
    out[7] = k2 * qyy;
    out[8] = k2 * qyz;
    out[9] = -k2 * (qxx + qyy);
    if (order == 2)
        return;
    const float qxxx = q[13], qxxy = q[14], qxyy = q[15], qyyy = q[16], qxxz = q[17],
                qxyz = q[18], qyyz = q[19];
    const float k3 = 5.0f * R * R * R;
    float *o = out + 10;
    // row x: (qxxx/2, qxyy/2, qxxy, qxxz, qxyz)            nln/pgrav.cu:139-153
    o[0] = k3 * 0.5f * qxxx; o[1] = k3 * 0.5f * qxyy; o[2] = k3 * qxxy; o[3] = k3 * qxxz; o[4] = k3 * qxyz;
    // row y: (qxxy/2, qyyy/2, qxyy, qxyz, qyyz)
    o[5] = k3 * 0.5f * qxxy; o[6] = k3 * 0.5f * qyyy; o[7] = k3 * qxyy; o[8] = k3 * qxyz; o[9] = k3 * qyyz;
    // row z: (qxxz/2, qyyz/2, qxyz, -(qxxx+qxyy), -(qxxy+qyyy))
    o[10] = k3 * 0.5f * qxxz; o[11] = k3 * 0.5f * qyyz; o[12] = k3 * qxyz;
    o[13] = -k3 * (qxxx + qxyy); o[14] = -k3 * (qxxy + qyyy);
    const float qxxxx = q[20], qxxxy = q[21], qxxyy = q[22], qxyyy = q[23], qyyyy = q[24],
                qxxxz = q[25], qxxyz = q[26], qxyyz = q[27], qyyyz = q[28];
    const float k4 = 26.25f * R * R * R * R;
    const float s6 = 1.0f / 6.0f;
    float *h = out + 25;
    // basis (U1..U7) = (6xxx, 6yyy, 2xxy, 6xxz, 2xyy, xyz, 6yyz) of nln/pgrav.cu:125-170
    h[0] = k4 * s6 * qxxxx; h[1] = k4 * s6 * qxyyy; h[2] = k4 * 0.5f * qxxxy; h[3] = k4 * s6 * qxxxz;
    h[4] = k4 * 0.5f * qxxyy; h[5] = k4 * qxxyz; h[6] = k4 * s6 * qxyyz;
    h[7] = k4 * s6 * qxxxy; h[8] = k4 * s6 * qyyyy; h[9] = k4 * 0.5f * qxxyy; h[10] = k4 * s6 * qxxyz;
    h[11] = k4 * 0.5f * qxyyy; h[12] = k4 * qxyyz; h[13] = k4 * s6 * qyyyz;
    h[14] = k4 * s6 * qxxxz; h[15] = k4 * s6 * qyyyz; h[16] = k4 * 0.5f * qxxyz;
    h[17] = -k4 * s6 * (qxxxx + qxxyy); h[18] = k4 * 0.5f * qxyyz; h[19] = -k4 * (qxyyy + qxxxy);
    h[20] = -k4 * s6 * (qyyyy + qxxyy);
}

// ---------------------------------------------------------------- interactions
// Order-4 (and order-2 when ORDER == 2) multipole interaction.  M points at a
// prepared record; the caller passes m separately (it may come from LDS).
// QSIGN = -1 reproduces nln/pgrav.cu's "phi -= eqe" for the quadrupole term.
template <int ORDER, bool DIPOLE, int QSIGN>
__device__ __forceinline__ void cell_interact(const float *__restrict__ M, float m, float x,
                                              float y, float z, float eps2, Acc4 &a) {
    const float x2 = x * x, y2 = y * y, z2 = z * z;
    const float r2 = x2 + y2 + z2 + eps2;
    const float rinv = -frsq(r2);
    const float ri2 = rinv * rinv;

    // order 1
    float g1x = 0.f, g1y = 0.f, g1z = 0.f, p1 = 0.f;
    if (DIPOLE) {
        g1x = M[0]; g1y = M[1]; g1z = M[2];
        p1 = g1x * x + g1y * y + g1z * z;
    }
    // order 2
    const float g2x = M[4] * x + M[5] * y + M[6] * z;
    const float g2y = M[5] * x + M[7] * y + M[8] * z;
    const float g2z = M[6] * x + M[8] * y + M[9] * z;
    const float p2 = g2x * x + g2y * y + g2z * z;
    const float rc = 2.0f * ri2;
    if (ORDER == 2) {
        float sphi = (QSIGN > 0 ? p2 : -p2);
        sphi = m + ri2 * (p1 + ri2 * sphi);
        const float tp = p1 + rc * p2;
        const float vx = g1x + rc * g2x, vy = g1y + rc * g2y, vz = g1z + rc * g2z;
        float sa = sphi + rc * tp;
        if (QSIGN < 0) sa += 2.0f * ri2 * ri2 * p2;
        const float ra = rinv * ri2;
        a.w += rinv * sphi;
        a.x += ra * (x * sa - vx);
        a.y += ra * (y * sa - vy);
        a.z += ra * (z * sa - vz);
        return;
    }
    // order 3: basis A, B, xy, xz, yz
    const float A = x2 - z2, B = y2 - z2;
    const float xy = x * y, xz = x * z, yz = y * z;
    const float *o = M + 10;
    const float g3x = o[0] * A + o[1] * B + o[2] * xy + o[3] * xz + o[4] * yz;
    const float g3y = o[5] * A + o[6] * B + o[7] * xy + o[8] * xz + o[9] * yz;
    const float g3z = o[10] * A + o[11] * B + o[12] * xy + o[13] * xz + o[14] * yz;
    const float p3 = g3x * x + g3y * y + g3z * z;
    // order 4: basis U1..U7
    const float z3 = 3.0f * z2;
    const float U1 = x * (x2 - z3);
    const float U2 = y * (y2 - z3);
    const float U3 = y * A;
    const float U4 = z * (3.0f * x2 - z2);
    const float U5 = x * B;
    const float U6 = xy * z;
    const float U7 = z * (3.0f * y2 - z2);
    const float *h = M + 25;
    const float g4x = h[0] * U1 + h[1] * U2 + h[2] * U3 + h[3] * U4 + h[4] * U5 + h[5] * U6 + h[6] * U7;
    const float g4y = h[7] * U1 + h[8] * U2 + h[9] * U3 + h[10] * U4 + h[11] * U5 + h[12] * U6 + h[13] * U7;
    const float g4z = h[14] * U1 + h[15] * U2 + h[16] * U3 + h[17] * U4 + h[18] * U5 + h[19] * U6 + h[20] * U7;
    const float p4 = g4x * x + g4y * y + g4z * z;

    const float rb = 1.5f * ri2, rq = (4.0f / 3.0f) * ri2;
    float sphi = p3 + ri2 * p4;
    sphi = (QSIGN > 0 ? p2 : -p2) + ri2 * sphi;
    sphi = m + ri2 * (p1 + ri2 * sphi);
    const float tp = p1 + rc * (p2 + rb * (p3 + rq * p4));
    const float vx = g1x + rc * (g2x + rb * (g3x + rq * g4x));
    const float vy = g1y + rc * (g2y + rb * (g3y + rq * g4y));
    const float vz = g1z + rc * (g2z + rb * (g3z + rq * g4z));
    float sa = sphi + rc * tp;
    if (QSIGN < 0) sa += 2.0f * ri2 * ri2 * p2;
    const float ra = rinv * ri2;
    a.w += rinv * sphi;
    a.x += ra * (x * sa - vx);
    a.y += ra * (y * sa - vy);
    a.z += ra * (z * sa - vz);
}

// Monopole / body with Plummer softening eps2 (0 = none), nln/pgrav.cu:380-400,
// nln/do_grav_sse4.c:2188-2215.
__device__ __forceinline__ void mono_interact(float m, float x, float y, float z, float eps2,
                                              Acc4 &a) {
    const float r2 = x * x + y * y + z * z + eps2;
    const float rinv = -frsq(r2);
    const float mr = m * rinv;
    const float mr3 = mr * rinv * rinv;
    a.w += mr;
    a.x += x * mr3;
    a.y += y * mr3;
    a.z += z * mr3;
}

// Body pair with Dehnen K1 softening (nln/pgrav.cu:745-770).  eps_inv = 1/eps.
__device__ __forceinline__ void k1_interact(float m, float x, float y, float z, float eps_inv,
                                            float eps2, Acc4 &a) {
    const float r2 = x * x + y * y + z * z;
    float phi, eqe;
    if (r2 > eps2) {
        const float rinv = -frsq(r2);
        phi = m * rinv;
        eqe = phi * rinv * rinv;
    } else {
        const float eps_inv2 = eps_inv * eps_inv;
        const float u2 = r2 * eps_inv2 - 1.0f;
        float t = (((45.0f / 32.0f) * u2 + (-3.0f / 8.0f)) * u2 + 0.5f) * u2 - 1.0f;
        phi = t * m * eps_inv;
        eqe = (((-135.0f / 16.0f) * u2 + 1.5f) * u2 - 1.0f) * (m * eps_inv * eps_inv2);
    }
    a.w += phi;
    a.x += x * eqe;
    a.y += y * eqe;
    a.z += z * eqe;
}

// Error-free transformation accumulate (Knuth 2Sum): (s, c) += v.
__device__ __forceinline__ void two_sum_add(float &s, float &c, float v) {
    const float t = s + v;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (v - bp);
    s = t;
    c += err;
}

}  // namespace pgrav
