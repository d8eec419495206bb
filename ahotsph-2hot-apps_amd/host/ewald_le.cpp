// Host side of the explicit-image periodic scheme: the far-image lattice
// coefficients c_4, c_6, c_8, the matrix W taking the global moments to the
// polynomial coefficients of the far field (csrc/pgrav_le.h states the
// expansion), and an FP64 CPU evaluation of that far field.
//
// The reference hard-codes the lattice sums for nimage = 1 and 2 as eight
// constants (nln/ewald_le2.c:347-365: ewx4 = sum n_x^4 / |n|^9 and so on) and
// expands the polynomials by hand (o4, o6, o8 :49-337).  Here the sums are
// computed for any nimage: the far images inside the cube |n|_inf <= R are
// summed directly and the rest of space is taken as the continuum integral of
// P_k(n_x/|n|) / |n|^(k+1) outside the cube [-R-1/2, R+1/2]^3, done by
// Gauss-Legendre quadrature over the cube's faces (the integrand is harmonic,
// so the lattice sum and the integral differ by O(R^-(k+1))).
#include <math.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "pgrav_host.h"
#include "pgrav_le.h"

namespace pgrav {

static double legendre(int k, double u) {
    double p0 = 1.0, p1 = u;
    if (k == 0)
        return p0;
    for (int l = 1; l < k; l++) {
        const double p2 = ((2 * l + 1) * u * p1 - l * p0) / (l + 1);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// nodes and weights of the m-point Gauss-Legendre rule on [-1, 1] (Newton on P_m)
static void gauss_legendre(int m, std::vector<double> &x, std::vector<double> &w) {
    x.resize(m);
    w.resize(m);
    for (int i = 0; i < m; i++) {
        double z = cos(M_PI * (i + 0.75) / (m + 0.5));
        double dp = 1.0;
        for (int it = 0; it < 100; it++) {
            const double p = legendre(m, z), pm = legendre(m - 1, z);
            dp = m * (z * p - pm) / (z * z - 1.0);
            const double dz = p / dp;
            z -= dz;
            if (fabs(dz) < 1e-16)
                break;
        }
        const double p = legendre(m, z), pm = legendre(m - 1, z);
        dp = m * (z * p - pm) / (z * z - 1.0);
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

double le_lattice_coeff(int k, int nimage) {
    if (k < 4 || nimage < 0)
        throw std::invalid_argument("le_lattice_coeff: degree >= 4, nimage >= 0");
    const int R = 40;
    double s = 0.0;
    for (int i = -R; i <= R; i++)
        for (int j = -R; j <= R; j++)
            for (int l = -R; l <= R; l++) {
                if (std::max(abs(i), std::max(abs(j), abs(l))) <= nimage)
                    continue;
                const double r2 = (double)i * i + (double)j * j + (double)l * l, r = sqrt(r2);
                s += legendre(k, i / r) / pow(r, k + 1);
            }
    // outside the cube of half-edge a: int dOmega P_k(w_x) r_c^(2-k) / (k-2), with
    // r_c the distance to the cube surface; on the face x = a the direction is
    // (a, y, z) / r and dOmega = a dy dz / r^3
    std::vector<double> gx, gw;
    gauss_legendre(64, gx, gw);
    double face_x = 0.0, face_o = 0.0;
    for (size_t p = 0; p < gx.size(); p++)
        for (size_t q = 0; q < gx.size(); q++) {
            const double y = gx[p], z = gx[q], r = sqrt(1.0 + y * y + z * z);
            const double wt = gw[p] * gw[q] * pow(r, -1.0 - k);
            face_x += wt * legendre(k, 1.0 / r);  // faces x = +-a (P_k even)
            face_o += wt * legendre(k, y / r);    // faces y = +-a and z = +-a
        }
    const double a = R + 0.5;
    return s + (2.0 * face_x + 4.0 * face_o) / (k - 2) * pow(a, 2.0 - k);
}

// coefficient of x^a y^b z^c in K_k (cubic-invariant harmonic, unit x^k term)
static double cubic_harmonic(int a, int b, int c) {
    if ((a | b | c) & 1)
        return 0.0;
    int e[3] = {a, b, c};
    std::sort(e, e + 3);  // e[0] <= e[1] <= e[2]
    const int k = a + b + c;
    // harmonic and invariant under axis permutations and reflections; solved
    // from Laplacian = 0 degree by degree
    if (k == 4)
        return e[2] == 4 ? 1.0 : (e[2] == 2 && e[1] == 2 ? -3.0 : 0.0);
    if (k == 6) {
        if (e[2] == 6)
            return 1.0;
        if (e[2] == 4 && e[1] == 2)
            return -7.5;
        if (e[0] == 2)
            return 90.0;
        return 0.0;
    }
    if (k == 8) {
        if (e[2] == 8)
            return 1.0;
        if (e[2] == 6 && e[1] == 2)
            return -14.0;
        if (e[2] == 4 && e[1] == 4)
            return 35.0;
        return 0.0;  // x^4 y^2 z^2
    }
    return 0.0;
}

static double binom(int n, int r) {
    double c = 1;
    for (int i = 0; i < r; i++) c = c * (n - i) / (i + 1);
    return c;
}

// x^a y^b z^c integrated over the unit cube [-1/2, 1/2]^3
static double unit_cube_moment(const int *e) {
    double v = 1.0;
    for (int a = 0; a < 3; a++) {
        if (e[a] & 1)
            return 0.0;
        v /= (double)(e[a] + 1) * pow(2.0, e[a]);
    }
    return v;
}

void le_matrix(int nimage, std::vector<double> &W, bool sb) {
    const int NM = LE_NMON;
    W.assign((size_t)NM * NM, 0.0);
    double ck[LE_DEG + 1] = {0};
    for (int k = 4; k <= LE_DEG; k += 2) ck[k] = le_lattice_coeff(k, nimage);
    // -c_k K_k(x - x_j): the x^beta coefficient collects kappa_alpha C(alpha, beta)
    // (-x_j)^(alpha - beta) over alpha = beta + gamma of degree k
    for (int bi = 0; bi < NM; bi++) {
        const int *be = LE_MON.e[bi];
        for (int gi = 0; gi < NM; gi++) {
            const int *ga = LE_MON.e[gi];
            const int al[3] = {be[0] + ga[0], be[1] + ga[1], be[2] + ga[2]};
            const int k = al[0] + al[1] + al[2];
            if (k < 4 || k > LE_DEG || (k & 1))
                continue;
            const double kap = cubic_harmonic(al[0], al[1], al[2]);
            if (kap == 0.0)
                continue;
            const double mult = binom(al[0], be[0]) * binom(al[1], be[1]) * binom(al[2], be[2]);
            const double sgn = ((ga[0] + ga[1] + ga[2]) & 1) ? -1.0 : 1.0;
            W[(size_t)bi * NM + gi] += -ck[k] * kap * mult * sgn;
        }
    }
    if (sb) {
        // the sources less the box's uniform cube of the total mass mu_0: mu_gamma ->
        // mu_gamma - mu_0 kappa_gamma (calculate_cartesian_moments with msb,
        // nln/ewald_le2.c:605-646; kappa_0 = 1 gives Q[0] = 0, :688-691), folded into
        // column 0 of W
        std::vector<double> wk(NM, 0.0);
        for (int bi = 0; bi < NM; bi++)
            for (int gi = 0; gi < NM; gi++)
                wk[bi] += W[(size_t)bi * NM + gi] * unit_cube_moment(LE_MON.e[gi]);
        for (int bi = 0; bi < NM; bi++) W[(size_t)bi * NM + 0] -= wk[bi];
        return;
    }
    // background (nln/ewald_le2.c:26-48): per unit source mass the constant
    // phicorr + xi and -(2 pi / 3) |x - x_j|^2; phicorr = sum over the explicit
    // images n != 0 of 1 / |n| (nln/main.c:1826-1834)
    double phicorr = 0.0;
    for (int i = -nimage; i <= nimage; i++)
        for (int j = -nimage; j <= nimage; j++)
            for (int l = -nimage; l <= nimage; l++)
                if (i || j || l)
                    phicorr += 1.0 / sqrt((double)(i * i + j * j + l * l));
    const double q = 2.0 * M_PI / 3.0;
    W[0] += phicorr + LE_XI;
    for (int ax = 0; ax < 3; ax++) {
        int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
        e1[ax] = 1;
        e2[ax] = 2;
        const int i1 = le_index(e1[0], e1[1], e1[2]), i2 = le_index(e2[0], e2[1], e2[2]);
        W[(size_t)0 * NM + i2] += -q;        // -(2 pi/3) sum_j m_j x_j^2
        W[(size_t)i1 * NM + i1] += 2.0 * q;  // +(4 pi/3) x . sum_j m_j x_j
        W[(size_t)i2 * NM + 0] += -q;        // -(2 pi/3) x^2 sum_j m_j
    }
}

void le_evaluate_host(const double *pos3, const double *mass, int64_t n, double box, int nimage,
                      double G, double *acc4, bool sb) {
    const int NM = LE_NMON;
    std::vector<double> W;
    le_matrix(nimage, W, sb);
    auto powers = [&](const double *p, double px[][LE_DEG + 1]) {
        for (int a = 0; a < 3; a++) {
            const double x = (p[a] - 0.5 * box) / box;
            px[a][0] = 1.0;
            for (int k = 1; k <= LE_DEG; k++) px[a][k] = px[a][k - 1] * x;
        }
    };
    std::vector<double> mu(NM, 0.0), C(NM, 0.0);
    for (int64_t j = 0; j < n; j++) {
        double px[3][LE_DEG + 1];
        powers(pos3 + 3 * j, px);
        for (int g = 0; g < NM; g++)
            mu[g] += mass[j] * px[0][LE_MON.e[g][0]] * px[1][LE_MON.e[g][1]] * px[2][LE_MON.e[g][2]];
    }
    for (int b = 0; b < NM; b++) {
        double s = 0.0;
        for (int g = 0; g < NM; g++) s += W[(size_t)b * NM + g] * mu[g];
        C[b] = s;
    }
    for (int64_t i = 0; i < n; i++) {
        double px[3][LE_DEG + 1];
        powers(pos3 + 3 * i, px);
        double phi = 0.0, gr[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < NM; b++) {
            const int *e = LE_MON.e[b];
            phi += C[b] * px[0][e[0]] * px[1][e[1]] * px[2][e[2]];
            if (e[0])
                gr[0] += C[b] * e[0] * px[0][e[0] - 1] * px[1][e[1]] * px[2][e[2]];
            if (e[1])
                gr[1] += C[b] * e[1] * px[0][e[0]] * px[1][e[1] - 1] * px[2][e[2]];
            if (e[2])
                gr[2] += C[b] * e[2] * px[0][e[0]] * px[1][e[1]] * px[2][e[2] - 1];
        }
        if (sb) {
            // a -= (4 pi / 3) dipole (nln/ewald_le2.c:371-375: Q[4], Q[7], Q[11] are
            // -4, 6, 6 times the dipole over the mass); the potential keeps no
            // self term (that sits in ewald_background, not called)
            for (int a = 0; a < 3; a++) {
                int e[3] = {0, 0, 0};
                e[a] = 1;
                gr[a] += 4.0 * M_PI / 3.0 * mu[le_index(e[0], e[1], e[2])];
            }
            for (int a = 0; a < 3; a++) acc4[4 * i + a] = -G / (box * box) * gr[a];
            acc4[4 * i + 3] = G / box * phi;
            continue;
        }
        for (int a = 0; a < 3; a++) acc4[4 * i + a] = -G / (box * box) * gr[a];
        acc4[4 * i + 3] = G / box * (phi - LE_XI * mass[i]);
    }
}

}  // namespace pgrav
