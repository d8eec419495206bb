// Error-tolerance critical radii of the tree cells (the reference's AREL_MAC,
// nln/cofm.c:276-565), computed in FP64 after the moments.
//
// For a cell of mass m, extent b and absolute moments B_k = sum m |x - c|^k about
// its centre of mass, the acceleration error of its expansion through order p at
// distance r is bounded by (Salmon & Warren)
//
//     E_p(r) = [(p + 2) B_{p+1} r^-(p+1) - (p + 1) B_{p+2} r^-(p+2)] / (r - b)^2,
//
// p = 1, 2, 4 for the listed orders 0, 2, 4 (the dipole about the centre of mass
// vanishes).  A critical radius is the largest root of E_p(r) = tolerance, cleared
// of denominators into a polynomial in r and found by Newton's method from a point
// above it (nln/cofm.c:577-624).  Three tolerances per order: absolute ptol; relative
// rel_tol m / r^2; floor rel_tol0 m / r^2; rcrit = max(min(abs, rel), rel0).  The odd
// moments are bounded: B3 = B2 sqrt(B2 / m) (order 0), B3 = b B2, B5 = b B4.
//
// Where the reference aborts (a zero tolerance makes its start point infinite; B2 = 0),
// this code decides: rel_tol = 0 or rel_tol0 = 0 switch that criterion off, and a cell
// with B2 = 0 (its bodies coincide: the monopole is exact) gets radius 0.
#include <math.h>

#include <algorithm>
#include <stdexcept>

#include "pgrav_host.h"

namespace pgrav {

namespace {

// nln/cofm.c:25-50: the tolerance scale sigma(m), indexed by floor(ln m), for stol_max > 1
const double kSigmaM[25] = {1.0,   1.0,   1.0,   1.0,   1.0,   1.0,   1.0,   1.0,   1.0,
                            1.5,   2.0,   3.0,   6.0,   10.88, 15.18, 21.34, 31.58, 46.73,
                            74.63, 120.3, 199.9, 353.1, 611.2, 1117,  1983};

// Largest root of sum_{i<=n} a[i] r^i by Newton's method from `start` (above the root),
// stopping when a step is below frac * start; the result carries that margin added, as
// the reference's radii do.  A search that does not settle in 30 steps returns start.
double root_above(const double *a, int n, double start, double frac) {
    const double tol = frac * start;
    double x = start;
    for (int it = 0; it < 30; it++) {
        double p = a[n], dp = 0.0;
        for (int k = n; k > 0; k--) {
            dp = dp * x + p;
            p = p * x + a[k - 1];
        }
        const double step = p / dp;
        x -= step;
        if (fabs(step) < tol)
            return x + tol;
    }
    return start + tol;
}

}  // namespace

double arel_ptol(double m, double bmax, const TreeParams &p, double bmax0) {
    if (p.stol_max > 1.0) {
        if (m <= 1.0)
            return p.errtol;
        if (m > 7.2e10)
            return p.stol_max * p.errtol;
        const double s = kSigmaM[std::min(24, std::max(0, (int)log(m)))];
        return std::min(s, p.stol_max) * p.errtol;
    }
    return p.errtol / (1.0 + p.ptol_boost * bmax / bmax0);
}

void arel_rcrit(double m, double b, double B2, double B4, double B6, int64_t count,
                const TreeParams &p, double bmax0, double out[3], const double *sb35) {
    const bool quad = p.quad_ncut > 0 && count >= p.quad_ncut;
    const bool hexa = p.hexa_ncut > 0 && count >= p.hexa_ncut;
    out[0] = 0.0;
    out[1] = quad ? 0.0 : NAN;
    out[2] = hexa ? 0.0 : NAN;
    if (!(B2 > 0.0) || !(m > 0.0))
        return;
    const double pt = arel_ptol(m, b, p, bmax0);
    const bool use_rel = p.frac_tol > 0.0, use_rel0 = p.frac_tol0 > 0.0;
    const double rt = p.frac_tol, r0 = p.frac_tol0;
    // the three running radii; each order starts from the previous order's
    double ab, rel = INFINITY, rel0 = 0.0;
    auto combine = [&]() { return std::max(std::min(ab, rel), rel0); };

    // order 0: E_1(r) = tol  <=>  tol r^3 (r - b)^2 - 3 B2 r + 2 B3 = 0
    const double B3 = B2 * sqrt(B2 / m);
    {
        const double h = 0.5 * b;
        const double a[6] = {2 * B3, -3 * B2, 0.0, pt * b * b, -2 * pt * b, pt};
        ab = root_above(a, 5, h + sqrt(h * h + sqrt(3.0 * B2 / pt)), 0.01);
    }
    // relative: tol = t m / r^2  <=>  t m r (r - b)^2 - 3 B2 r + 2 B3 = 0
    auto rel_poly0 = [&](double t, double *a) {
        a[0] = 2 * B3;
        a[1] = -3 * B2 + t * b * b * m;
        a[2] = -2 * t * b * m;
        a[3] = t * m;
    };
    if (use_rel0) {
        double a[4];
        rel_poly0(r0, a);
        rel0 = root_above(a, 3, b + sqrt(3.0 / r0 * B2 / m), 0.01);
    }
    if (use_rel) {
        double a[4];
        rel_poly0(rt, a);
        rel = root_above(a, 3, b + sqrt(3.0 / rt * B2 / m), 0.01);
    }
    out[0] = combine();

    // order 2: E_2 with B3 <= b B2
    if (quad) {
        const double B3q = sb35 ? sb35[0] : b * B2;
        if (sb35)
            B4 = 0.0;
        const double a[7] = {3 * B4, -4 * B3q, 0.0, 0.0, pt * b * b, -2 * pt * b, pt};
        ab = root_above(a, 6, ab, 0.01);
        auto relq = [&](double t, double start) {
            const double c[5] = {3 * B4, -4 * B3q, t * b * b * m, -2 * t * b * m, t * m};
            return root_above(c, 4, start, 0.01);
        };
        if (use_rel0)
            rel0 = relq(r0, rel0);
        if (use_rel)
            rel = relq(rt, rel);
        out[1] = combine();
    }
    // order 4: E_4 with B5 <= b B4
    if (hexa) {
        const double B5 = sb35 ? sb35[1] : b * B4;
        if (sb35)
            B6 = 0.0;
        const double a[9] = {5 * B6, -6 * B5, 0.0, 0.0, 0.0, 0.0, pt * b * b, -2 * pt * b, pt};
        ab = root_above(a, 8, ab, 0.001);
        auto relh = [&](double t, double start) {
            const double c[7] = {5 * B6, -6 * B5, 0.0, 0.0, t * b * b * m, -2 * t * b * m, t * m};
            return root_above(c, 6, start, 0.001);
        };
        if (use_rel0)
            rel0 = relh(r0, rel0);
        if (use_rel)
            rel = relh(rt, rel);
        out[2] = combine();
    }
}

// B2 and B4 from the raw moments about the centre of mass (the sums the reference
// forms from its cofmdata, nln/cofm.c:306, :363, :438); B6 = sum m r^6 from the bodies.
void compute_rcrit(Tree &t, const TreeParams &p, int nth) {
    if (!(p.errtol > 0.0) || p.frac_tol < 0.0 || p.frac_tol0 < 0.0)
        throw std::runtime_error("the tolerance MAC needs errtol > 0 and frac_tol, frac_tol0 >= 0");
    const int64_t nn = t.nnodes();
    t.rcrit.assign(3 * (size_t)nn, 0.0);
    const int i200 = raw_index(2, 0, 0), i020 = raw_index(0, 2, 0), i002 = raw_index(0, 0, 2);
    const int i400 = raw_index(4, 0, 0), i040 = raw_index(0, 4, 0), i004 = raw_index(0, 0, 4);
    const int i220 = raw_index(2, 2, 0), i202 = raw_index(2, 0, 2), i022 = raw_index(0, 2, 2);
    const double bmax0 = 0.5 * t.size;  // the reference's r0: the system's half-extent
    parallel_for(nth, nn, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; c++) {
            const double *raw = &t.raw[35 * (size_t)c];
            const double m = raw[0];
            const double B2 = raw[i200] + raw[i020] + raw[i002];
            const double B4 = raw[i400] + raw[i040] + raw[i004] + 2 * raw[i220] + 2 * raw[i202] +
                              2 * raw[i022];
            double B6 = 0.0;
            const double *cm = &t.cofm[3 * c];
            for (int64_t i = t.first[c]; i < t.first[c] + t.count[c]; i++) {
                const double dx = t.pos[3 * i] - cm[0], dy = t.pos[3 * i + 1] - cm[1],
                             dz = t.pos[3 * i + 2] - cm[2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                B6 += t.mass[i] * r2 * r2 * r2;
            }
            arel_rcrit(m, t.bmax[c], B2, B4, B6, t.count[c], p, bmax0, &t.rcrit[3 * (size_t)c]);
        }
    });
}

}  // namespace pgrav
