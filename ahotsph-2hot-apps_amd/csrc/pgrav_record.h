// Source-record preparation shared by host (FP64 tables) and device (drop-in
// path).  No HIP dependency: the same template is compiled for both sides.
//
// A reference record (nln/pgrav.cu:16-47 field order: mass, x, y, z, R, qx, qy,
// qz, 5 quadrupole, 7 octupole, 9 hexadecapole components, normalised by R^k as
// nln/cofm.c:398-562 stores them) becomes the prepared layout documented in
// pgrav_device.h.
#pragma once

#if defined(__HIPCC__)
#define PGRAV_HD __host__ __device__
#else
#define PGRAV_HD
#endif

#define PREC4 48
#define PREC2 12
#define PREC_TRACE 46  // prepared-record slot of sum m |x - cofm|^2 (periodic potentials)

namespace pgrav {

template <typename T>
PGRAV_HD inline void prep_record(const T *q, int order, int dipole, T *out /* PREC4 */) {
    for (int k = 0; k < PREC4; k++) out[k] = T(0);
    out[3] = q[0];
    if (order == 0)
        return;
    const T R = q[4];
    if (dipole) {
        out[0] = R * q[5];
        out[1] = R * q[6];
        out[2] = R * q[7];
    }
    const T qxx = q[8], qxy = q[9], qyy = q[10], qxz = q[11], qyz = q[12];
    const T k2 = T(1.5) * R * R;
    out[4] = k2 * qxx;
    out[5] = k2 * qxy;
    out[6] = k2 * qxz;
    out[7] = k2 * qyy;
    out[8] = k2 * qyz;
    out[9] = -k2 * (qxx + qyy);
    if (order == 2)
        return;
    const T qxxx = q[13], qxxy = q[14], qxyy = q[15], qyyy = q[16], qxxz = q[17], qxyz = q[18],
            qyyz = q[19];
    const T k3 = T(5) * R * R * R;
    const T h2 = T(0.5);
    T *o = out + 10;
    // row x: (qxxx/2, qxyy/2, qxxy, qxxz, qxyz)            nln/pgrav.cu:139-153
    o[0] = k3 * h2 * qxxx; o[1] = k3 * h2 * qxyy; o[2] = k3 * qxxy; o[3] = k3 * qxxz; o[4] = k3 * qxyz;
    // row y: (qxxy/2, qyyy/2, qxyy, qxyz, qyyz)
    o[5] = k3 * h2 * qxxy; o[6] = k3 * h2 * qyyy; o[7] = k3 * qxyy; o[8] = k3 * qxyz; o[9] = k3 * qyyz;
    // row z: (qxxz/2, qyyz/2, qxyz, -(qxxx+qxyy), -(qxxy+qyyy))
    o[10] = k3 * h2 * qxxz; o[11] = k3 * h2 * qyyz; o[12] = k3 * qxyz;
    o[13] = -k3 * (qxxx + qxyy); o[14] = -k3 * (qxxy + qyyy);
    const T qxxxx = q[20], qxxxy = q[21], qxxyy = q[22], qxyyy = q[23], qyyyy = q[24],
            qxxxz = q[25], qxxyz = q[26], qxyyz = q[27], qyyyz = q[28];
    const T k4 = T(26.25) * R * R * R * R;
    const T s6 = T(1) / T(6);
    T *h = out + 25;
    // basis (U1..U7) = (6xxx, 6yyy, 2xxy, 6xxz, 2xyy, xyz, 6yyz) of nln/pgrav.cu:125-170
    h[0] = k4 * s6 * qxxxx; h[1] = k4 * s6 * qxyyy; h[2] = k4 * h2 * qxxxy; h[3] = k4 * s6 * qxxxz;
    h[4] = k4 * h2 * qxxyy; h[5] = k4 * qxxyz; h[6] = k4 * s6 * qxyyz;
    h[7] = k4 * s6 * qxxxy; h[8] = k4 * s6 * qyyyy; h[9] = k4 * h2 * qxxyy; h[10] = k4 * s6 * qxxyz;
    h[11] = k4 * h2 * qxyyy; h[12] = k4 * qxyyz; h[13] = k4 * s6 * qyyyz;
    h[14] = k4 * s6 * qxxxz; h[15] = k4 * s6 * qyyyz; h[16] = k4 * h2 * qxxyz;
    h[17] = -k4 * s6 * (qxxxx + qxxyy); h[18] = k4 * h2 * qxyyz; h[19] = -k4 * (qxyyy + qxxxy);
    h[20] = -k4 * s6 * (qyyyy + qxxyy);
}

// Trace of the raw second moments, sum m |x - cofm|^2: the periodic kernels add
// the cells' Ewald correction at the centre of mass only, and the second-order
// remainder of its potential is -(2 pi / 3) / L^3 times this trace (the trace of
// the correction's Hessian is -4 pi / L^3 everywhere; the traceless rest of the
// cell's second moments couples to the Hessian's small traceless part).
template <typename T>
PGRAV_HD inline T raw_trace(const T *u) {
    return u[4] + u[7] + u[9];
}

// Raw moments about the cell centre (35 values in the order of
// nln/cofm.c:660-720: m; x y z; x2 xy xz y2 yz z2; x3 x2y x2z xy2 xyz xz2 y3 y2z
// yz2 z3; x4 x3y x3z x2y2 x2yz x2z2 xy3 xy2z xyz2 xz3 y4 y3z y2z2 yz3 z4) ->
// the 29-field reference record with R = bmax (nln/cofm.c:398-562).
template <typename T>
PGRAV_HD inline void raw_to_record(const T *u, const T *center, T bmax, T *rec /* 29 */) {
    enum { M, X, Y, Z, X2, XY, XZ, Y2, YZ, Z2, X3, X2Y, X2Z, XY2, XYZ, XZ2, Y3, Y2Z, YZ2, Z3,
           X4, X3Y, X3Z, X2Y2, X2YZ, X2Z2, XY3, XY2Z, XYZ2, XZ3, Y4, Y3Z, Y2Z2, YZ3, Z4 };
    const T rinv = T(1) / bmax;
    rec[0] = u[M];
    rec[1] = center[0];
    rec[2] = center[1];
    rec[3] = center[2];
    rec[4] = bmax;
    rec[5] = rinv * u[X];
    rec[6] = rinv * u[Y];
    rec[7] = rinv * u[Z];
    T t = (u[X2] + u[Y2] + u[Z2]) / T(3);
    const T r2 = rinv * rinv;
    rec[8] = (u[X2] - t) * r2;
    rec[9] = u[XY] * r2;
    rec[10] = (u[Y2] - t) * r2;
    rec[11] = u[XZ] * r2;
    rec[12] = u[YZ] * r2;
    const T tx = (u[X3] + u[XY2] + u[XZ2]) / T(5);
    const T ty = (u[X2Y] + u[Y3] + u[YZ2]) / T(5);
    const T tz = (u[X2Z] + u[Y2Z] + u[Z3]) / T(5);
    const T r3 = r2 * rinv;
    rec[13] = (u[X3] - T(3) * tx) * r3;
    rec[14] = (u[X2Y] - ty) * r3;
    rec[15] = (u[XY2] - tx) * r3;
    rec[16] = (u[Y3] - T(3) * ty) * r3;
    rec[17] = (u[X2Z] - tz) * r3;
    rec[18] = u[XYZ] * r3;
    rec[19] = (u[Y2Z] - tz) * r3;
    const T txx = (u[X4] + u[X2Y2] + u[X2Z2]) / T(7);
    const T txy = (u[X3Y] + u[XY3] + u[XYZ2]) / T(7);
    const T txz = (u[X3Z] + u[XY2Z] + u[XZ3]) / T(7);
    const T tyy = (u[X2Y2] + u[Y4] + u[Y2Z2]) / T(7);
    const T tyz = (u[X2YZ] + u[Y3Z] + u[YZ3]) / T(7);
    const T tzz = (u[X2Z2] + u[Y2Z2] + u[Z4]) / T(7);
    t = T(0.1) * (txx + tyy + tzz);
    const T r4 = r2 * r2;
    rec[20] = (u[X4] - T(6) * (txx - t)) * r4;
    rec[21] = (u[X3Y] - T(3) * txy) * r4;
    rec[22] = (u[X2Y2] - (txx + tyy - T(2) * t)) * r4;
    rec[23] = (u[XY3] - T(3) * txy) * r4;
    rec[24] = (u[Y4] - T(6) * (tyy - t)) * r4;
    rec[25] = (u[X3Z] - T(3) * txz) * r4;
    rec[26] = (u[X2YZ] - tyz) * r4;
    rec[27] = (u[XY2Z] - txz) * r4;
    rec[28] = (u[Y3Z] - T(3) * tyz) * r4;
}

}  // namespace pgrav
