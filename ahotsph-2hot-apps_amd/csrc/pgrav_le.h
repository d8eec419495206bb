// Periodic far field of the explicit-image scheme: the local expansion (LE)
// of the lattice images beyond the (2n+1)^3 that the walk lists explicitly,
// plus the uniform-background terms (nln/ewald_le2.c:26-48, :339-381 and
// nln/main.c:1811-1862, FixCubeEwaldLE).  Shared by the host (coefficient
// matrix, CPU evaluation) and the device kernels (pgrav_ewald_le.hip).
//
// In box units (positions x = (p - L/2) / L, the box centred on the origin):
//
//   F(r) = sum over far images n of 1/|r - n| = const + (2 pi / 3) r^2
//          + sum_{k = 4, 6, 8} c_k K_k(r) + O(r^10),
//
// where K_k is the cubic-invariant harmonic polynomial of degree k with unit
// x^k coefficient (cubic symmetry leaves one per degree; odd degrees and the
// degree-2 harmonic vanish) and c_k = sum_far P_k(n_x / |n|) / |n|^(k+1), the
// x^k coefficient of the Taylor series of 1/|r - n| summed over the far
// images (P_k the Legendre polynomial).  With phi = -G m / r per source, the
// far images of all bodies add, at x,
//
//   phi(x) = (G / L) [ sum_beta C_beta x^beta - xi m_i ]
//   C      = W mu,   mu_gamma = sum_j m_j x_j^gamma  (global moments, |gamma| <= 8)
//
// and a = -(G / L^2) grad_x sum_beta C_beta x^beta.  W holds the multinomial
// expansion of -c_k K_k(x - x_j) and the background terms of nln/ewald_le2.c:
// 26-48 (phicorr + xi - (2 pi / 3) |x - x_j|^2 per unit mass; xi the lattice
// self constant, 2 x 1.4186487533 in nln/ewald_le2.c:38-43).
#pragma once

namespace pgrav {

#define LE_DEG 8     // highest expansion degree (orders 4, 6, 8: nln/ewald_le2.c:365-367)
#define LE_NMON 165  // monomials x^a y^b z^c with a + b + c <= 8

// Monomials of total degree <= LE_DEG: by degree, then a descending, then b
// descending.
struct LeMonomials {
    int e[LE_NMON][3];
};

constexpr LeMonomials le_make_monomials() {
    LeMonomials m{};
    int k = 0;
    for (int d = 0; d <= LE_DEG; d++)
        for (int a = d; a >= 0; a--)
            for (int b = d - a; b >= 0; b--) {
                m.e[k][0] = a;
                m.e[k][1] = b;
                m.e[k][2] = d - a - b;
                k++;
            }
    return m;
}

constexpr LeMonomials LE_MON = le_make_monomials();

constexpr int le_index(int a, int b, int c) {
    const int d = a + b + c;
    return d * (d + 1) * (d + 2) / 6 + (d - a) * (d - a + 1) / 2 + (d - a - b);
}

static_assert(le_index(0, 0, 0) == 0 && le_index(1, 0, 0) == 1 && le_index(0, 0, 8) == LE_NMON - 1,
              "monomial index");

// The lattice self constant xi: lim_{r->0} (psi(r) - 1/r) = -xi for the Ewald
// potential psi of a unit cubic lattice with a uniform neutralising background.
constexpr double LE_XI = 2.837297479480620;

}  // namespace pgrav
