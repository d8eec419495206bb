// The Ewald correction of the nearest-image periodic mode (periodic = 1).
//
// For a unit source at r = source - sink (the nearest image, |r_i| <= L/2) the
// pull of all its other periodic images and of the neutralising background is
// (C(r), psi(r)): C = f_ewald(r) - r/|r|^3 and psi = phi_ewald(r) + 1/|r|, with
// f_ewald and phi_ewald the lattice sums of nln/ewald.c:13-64.  C = grad psi is
// smooth inside the cube; its Jacobian J (symmetric, trace -4 pi / L^3) and
// second derivatives T (fully symmetric, traceless in any pair) are tabulated
// with it.
//
// Table: the nodes (i h, j h, k h), h = (L/2)/(EWALD_N - 1), of the positive
// octant with i >= j >= k only -- the field has the cube's symmetry, so the
// other orderings and octants follow by permuting and reflecting axes --
// numbered i(i+1)(i+2)/6 + j(j+1)/2 + k (45,760 nodes).  Each node holds 20
// values, in the node's own (sorted) frame:
//   [0] psi  [1..3] C  [4..9] J (xx xy xz yy yz zz)
//   [10..19] T (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
// FP64 (k_ewald_table, closed form) and an FP32 copy: 7.3 MB and 3.7 MB.
//
// A lookup at r sorts |r_i| with a 3-element network (remembering the swaps),
// takes the nearest node, steps from it by Taylor's formula to third order in
// psi, second in C and first in J, then undoes the swaps and reflections.
#pragma once
#include <hip/hip_runtime.h>

#include "pgrav_device.h"

#define EWALD_N 64  // table points per axis over [0, L/2]
#define EW_REC 20   // values per node
#define EWALD_NODES (EWALD_N * (EWALD_N + 1) * (EWALD_N + 2) / 6)

namespace pgrav {

__host__ __device__ __forceinline__ int ew_node(int i, int j, int k) {
    return i * (i + 1) * (i + 2) / 6 + j * (j + 1) / 2 + k;
}

// psi, C, J and T at one point (J and T in the symmetric orders above)
template <typename T>
struct EwField {
    T psi, c[3], j[6], t[10];
};

template <typename T>
__device__ __forceinline__ void swap_if(bool b, T &x, T &y) {
    const T a = b ? y : x, c = b ? x : y;
    x = a;
    y = c;
}

// transposition of axes 0 and 1 (x <-> y) of a field
template <typename T>
__device__ __forceinline__ void transpose01(bool b, EwField<T> &F) {
    swap_if(b, F.c[0], F.c[1]);
    swap_if(b, F.j[0], F.j[3]);  // xx yy
    swap_if(b, F.j[2], F.j[4]);  // xz yz
    swap_if(b, F.t[0], F.t[6]);  // xxx yyy
    swap_if(b, F.t[1], F.t[3]);  // xxy xyy
    swap_if(b, F.t[2], F.t[7]);  // xxz yyz
    swap_if(b, F.t[5], F.t[8]);  // xzz yzz
}

// transposition of axes 1 and 2 (y <-> z)
template <typename T>
__device__ __forceinline__ void transpose12(bool b, EwField<T> &F) {
    swap_if(b, F.c[1], F.c[2]);
    swap_if(b, F.j[3], F.j[5]);  // yy zz
    swap_if(b, F.j[1], F.j[2]);  // xy xz
    swap_if(b, F.t[6], F.t[9]);  // yyy zzz
    swap_if(b, F.t[1], F.t[2]);  // xxy xxz
    swap_if(b, F.t[3], F.t[5]);  // xyy xzz
    swap_if(b, F.t[7], F.t[8]);  // yyz yzz
}

// The node nearest (|r| sorted) and the sorting network's swaps.
template <typename T>
struct EwNode {
    int node;
    T d0, d1, d2;       // offset from the node, sorted frame
    bool w1, w2, w3;    // swaps (0,1), (1,2), (0,1) that sorted |r|
    T s0, s1, s2;       // signs of r
};

template <typename T>
__device__ __forceinline__ EwNode<T> ew_locate(T rx, T ry, T rz, T box) {
    EwNode<T> n;
    n.s0 = rx < T(0) ? T(-1) : T(1);
    n.s1 = ry < T(0) ? T(-1) : T(1);
    n.s2 = rz < T(0) ? T(-1) : T(1);
    T a0 = fabs(rx), a1 = fabs(ry), a2 = fabs(rz);
    n.w1 = a0 < a1;
    swap_if(n.w1, a0, a1);
    n.w2 = a1 < a2;
    swap_if(n.w2, a1, a2);
    n.w3 = a0 < a1;
    swap_if(n.w3, a0, a1);
    const T h = T(0.5) * box / T(EWALD_N - 1);
    const T ih = T(1) / h;
    // clamped both ways: a NaN offset must not become an out-of-range index
    const int i = max(0, min((int)rint(a0 * ih), EWALD_N - 1));
    const int j = max(0, min((int)rint(a1 * ih), i));
    const int k = max(0, min((int)rint(a2 * ih), j));
    n.node = ew_node(i, j, k);
    n.d0 = a0 - T(i) * h;
    n.d1 = a1 - T(j) * h;
    n.d2 = a2 - T(k) * h;
    return n;
}

template <typename TV>
__device__ __forceinline__ void ew_load(const TV *__restrict__ tab, int node, TV *v) {
#pragma unroll
    for (int q = 0; q < EW_REC; q++) v[q] = tab[(size_t)node * EW_REC + q];
}

// Taylor step of the node's record by d in the sorted frame: psi to third
// order, C to second, J to first; T is the node's.
template <typename T, typename TV>
__device__ __forceinline__ void ew_taylor(const TV *v, T d0, T d1, T d2, EwField<T> &F) {
    const T t[10] = {T(v[10]), T(v[11]), T(v[12]), T(v[13]), T(v[14]),
                     T(v[15]), T(v[16]), T(v[17]), T(v[18]), T(v[19])};
    // (T d)_ab, symmetric: xx xy xz yy yz zz
    const T tdxx = t[0] * d0 + t[1] * d1 + t[2] * d2;
    const T tdxy = t[1] * d0 + t[3] * d1 + t[4] * d2;
    const T tdxz = t[2] * d0 + t[4] * d1 + t[5] * d2;
    const T tdyy = t[3] * d0 + t[6] * d1 + t[7] * d2;
    const T tdyz = t[4] * d0 + t[7] * d1 + t[8] * d2;
    const T tdzz = t[5] * d0 + t[8] * d1 + t[9] * d2;
    const T jxx = T(v[4]) + tdxx, jxy = T(v[5]) + tdxy, jxz = T(v[6]) + tdxz;
    const T jyy = T(v[7]) + tdyy, jyz = T(v[8]) + tdyz, jzz = T(v[9]) + tdzz;
    // J(d) = J + T d; C(d) = C + J d + T d d / 2 = C + (J + J(d)) d / 2
    const T h = T(0.5);
    const T cx = T(v[1]) + h * ((T(v[4]) + jxx) * d0 + (T(v[5]) + jxy) * d1 + (T(v[6]) + jxz) * d2);
    const T cy = T(v[2]) + h * ((T(v[5]) + jxy) * d0 + (T(v[7]) + jyy) * d1 + (T(v[8]) + jyz) * d2);
    const T cz = T(v[3]) + h * ((T(v[6]) + jxz) * d0 + (T(v[8]) + jyz) * d1 + (T(v[9]) + jzz) * d2);
    // psi(d) = psi + C.d + d.J.d / 2 + T d d d / 6 = psi + (2 C + C(d)) . d / 3 + d.J.d / 6
    // (C(d) . d = C.d + d.J.d + T d d d / 2)
    const T c0d = T(v[1]) * d0 + T(v[2]) * d1 + T(v[3]) * d2;
    const T cdd = cx * d0 + cy * d1 + cz * d2;
    const T djd = d0 * (T(v[4]) * d0 + T(2) * (T(v[5]) * d1 + T(v[6]) * d2)) +
                  d1 * (T(v[7]) * d1 + T(2) * T(v[8]) * d2) + d2 * T(v[9]) * d2;
    F.psi = T(v[0]) + (T(2) * c0d + cdd) * T(1.0 / 3.0) + djd * T(1.0 / 6.0);
    F.c[0] = cx;
    F.c[1] = cy;
    F.c[2] = cz;
    F.j[0] = jxx;
    F.j[1] = jxy;
    F.j[2] = jxz;
    F.j[3] = jyy;
    F.j[4] = jyz;
    F.j[5] = jzz;
#pragma unroll
    for (int q = 0; q < 10; q++) F.t[q] = t[q];
}

// The field at r = (rx, ry, rz), |r_i| <= L/2.
template <typename T, typename TV>
__device__ __forceinline__ void ewald_field(const TV *__restrict__ tab, T rx, T ry, T rz, T box,
                                            EwField<T> &F) {
    const EwNode<T> n = ew_locate<T>(rx, ry, rz, box);
    TV v[EW_REC];
    ew_load(tab, n.node, v);
    ew_taylor<T>(v, n.d0, n.d1, n.d2, F);
    // back from the sorted frame (the swaps undone in reverse order) ...
    transpose01(n.w3, F);
    transpose12(n.w2, F);
    transpose01(n.w1, F);
    // ... and from the positive octant: C odd in each axis, J_ab with s_a s_b,
    // T_abc with s_a s_b s_c, psi even
    const T s0 = n.s0, s1 = n.s1, s2 = n.s2;
    const T s01 = s0 * s1, s02 = s0 * s2, s12 = s1 * s2, s012 = s01 * s2;
    F.c[0] *= s0;
    F.c[1] *= s1;
    F.c[2] *= s2;
    F.j[1] *= s01;
    F.j[2] *= s02;
    F.j[4] *= s12;
    F.t[0] *= s0;   // xxx
    F.t[1] *= s1;   // xxy
    F.t[2] *= s2;   // xxz
    F.t[3] *= s0;   // xyy
    F.t[4] *= s012; // xyz
    F.t[5] *= s0;   // xzz
    F.t[6] *= s1;   // yyy
    F.t[7] *= s2;   // yyz
    F.t[8] *= s1;   // yzz
    F.t[9] *= s2;   // zzz
}

// One pair's correction: x = sink - source (the nearest image), so r = -x: the
// sink takes m C(r) and m psi(r).  Only C and psi are formed.
template <typename T, typename TV>
__device__ __forceinline__ void ewald_pair(const TV *__restrict__ tab, T m, T x, T y, T z, T box,
                                           Acc4T<T> &a) {
    EwField<T> F;
    ewald_field<T, TV>(tab, -x, -y, -z, box, F);
    a.x += m * F.c[0];
    a.y += m * F.c[1];
    a.z += m * F.c[2];
    a.w += m * F.psi;
}

}  // namespace pgrav
