"""TEST INFRASTRUCTURE ONLY -- pure-Python restatement of the reference's
error-tolerance multipole acceptance (AREL_MAC).  Only tests/ import it, as the
checker.

Critical radii (nln/cofm.c:276-565, the AREL_MAC branch of CellFromCofm): for a
cell of mass m, extent b = bmax and absolute moments B_k = sum_j m_j |x_j - c|^k
about its centre of mass, the Salmon-Warren bound on the acceleration error of
its order-p expansion at distance r is

    E_p(r) = [(p + 2) B_{p+1} / r^(p+1) - (p + 1) B_{p+2} / r^(p+2)] / (r - b)^2

for the expansion through order p: p = 1 for the monopole about the centre of
mass (whose dipole vanishes), 2 for order 2, 4 for order 4.  The critical radius of each order is
the largest root of E_p(r) = tol, found by Newton's method from an upper
starting point (rtnewt, :595-624, with the Horner evaluation rcrit_poly,
:577-593), plus the step tolerance:

    abs   tol = ptol (the absolute tolerance)
    rel   tol = rel_tol  * m / r^2   (relative to the cell's own pull)
    rel0  tol = rel_tol0 * m / r^2

    rcrit = max(min(abs, rel), rel0)

with the odd moments replaced by bounds: B3 = B2 sqrt(B2 / m) for the monopole
(:310), B3 = b B2 (:361) and B5 = b B4 (:436) for the higher orders.  Order 2
starts its iterations from order 0's radii, order 4 from order 2's (the
variables carry over, :369-393 and :453-475).

Acceptance (nln/mac.c:1776-1800, DLRcritMAC): a source cell is taken at
order 0 if r^2 > (rcrit + b)^2, else at order 2 if it has quadrupole moments
(at least quad_ncut bodies) and r^2 > (rcrit_q + b)^2, else at order 4
(at least hexa_ncut bodies) if r^2 > (rcrit_h + b)^2; otherwise it is opened.

Decided deviations (the reference aborts in these cases): rel_tol = 0 turns
the relative criterion off (rel = +inf) and rel_tol0 = 0 the floor (rel0 = 0);
B2 = 0 (a single body, or coincident bodies: the monopole is exact) gives
radius 0.  parity pins: tests/test_mac_cpu.py.
"""
import math

import numpy as np

# nln/cofm.c:25-50: sigma(m) by floor(ln m), used when stol_max > 1
SIGMA_M = [1.0] * 9 + [1.5, 2.0, 3.0, 6.0, 10.88, 15.18, 21.34, 31.58, 46.73, 74.63, 120.3,
                       199.9, 353.1, 611.2, 1117.0, 1983.0]
MIN_SIGMA_M, MAX_SIGMA_M = 1.0, 7.2e10


def horner(a, n, r):
    """(p(r), p'(r)) of p = sum_{i<=n} a[i] r^i (nln/cofm.c:577-593)."""
    p, dp = a[n], 0.0
    while n > 0:
        dp = dp * r + p
        n -= 1
        p = p * r + a[n]
    return p, dp


def newton(a, n, x1, xacc, jmax=30):
    """nln/cofm.c:601-624: Newton from x1; returns x1 when 30 steps do not converge."""
    x = x1
    for _ in range(jmax):
        f, df = horner(a, n, x)
        dx = f / df
        x -= dx
        if abs(dx) < xacc:
            return x
    return x1


def ptol_of(m, bmax, tol, ptol_boost=0.0, bmax0=1.0, stol_max=0.0):
    """nln/cofm.c:285-304."""
    if stol_max > 1.0:
        if m <= MIN_SIGMA_M:
            return tol
        if m > MAX_SIGMA_M:
            return stol_max * tol
        s = SIGMA_M[min(24, int(math.log(m)))]  # ln(7.2e10) is 25.0: the last entry
        return s * tol if s < stol_max else stol_max * tol
    return tol / (1.0 + ptol_boost * bmax / bmax0)


def rcrit(m, bmax, B2, B4, B6, count, tol, rel_tol=0.0, rel_tol0=0.0, ptol_boost=0.0, bmax0=1.0,
          stol_max=0.0, quad_ncut=7, hexa_ncut=20):
    """(rcrit, rcrit_q, rcrit_h) of one cell (rcrit_q / rcrit_h: nan where the
    cell has no such moments)."""
    b = bmax
    if B2 <= 0.0:
        return (0.0, 0.0 if quad_ncut and count >= quad_ncut else math.nan,
                0.0 if hexa_ncut and count >= hexa_ncut else math.nan)
    ptol = ptol_of(m, b, tol, ptol_boost, bmax0, stol_max)

    def solve(a, n, start, frac):
        return newton(a, n, start, frac * start) + frac * start

    # order 0 (:305-345)
    half = 0.5 * b
    start_abs = half + math.sqrt(half * half + math.sqrt(3.0 * B2 / ptol))
    B3 = B2 * math.sqrt(B2 / m)
    a = [2 * B3, -3 * B2, 0.0, ptol * b * b, -2 * ptol * b, ptol]
    abs_r = solve(a, 5, start_abs, 0.01)
    rel0 = rel = None
    if rel_tol0 > 0:
        start_rel0 = b + math.sqrt(3.0 / rel_tol0 * B2 / m)
        a = [2 * B3, -3 * B2 + rel_tol0 * b * b * m, -2 * rel_tol0 * b * m, rel_tol0 * m]
        rel0 = solve(a, 3, start_rel0, 0.01)
    if rel_tol > 0:
        start_rel = b + math.sqrt(3.0 / rel_tol * B2 / m)
        a = [2 * B3, -3 * B2 + rel_tol * b * b * m, -2 * rel_tol * b * m, rel_tol * m]
        rel = solve(a, 3, start_rel, 0.01)

    def combine(ab, rl, r0):
        r = ab if rl is None else min(ab, rl)
        return r0 if r0 is not None and r0 > r else r

    out0 = combine(abs_r, rel, rel0)
    outq = outh = math.nan
    # order 2 (:361-396), no background subtraction
    if quad_ncut and count >= quad_ncut:
        B3q = b * B2
        a = [3 * B4, -4 * B3q, 0.0, 0.0, ptol * b * b, -2 * ptol * b, ptol]
        abs_r = solve(a, 6, abs_r, 0.01)
        if rel0 is not None:
            a = [3 * B4, -4 * B3q, rel_tol0 * b * b * m, -2 * rel_tol0 * b * m, rel_tol0 * m]
            rel0 = solve(a, 4, rel0, 0.01)
        if rel is not None:
            a = [3 * B4, -4 * B3q, rel_tol * b * b * m, -2 * rel_tol * b * m, rel_tol * m]
            rel = solve(a, 4, rel, 0.01)
        outq = combine(abs_r, rel, rel0)
    # order 4 (:436-482)
    if hexa_ncut and count >= hexa_ncut:
        B5 = b * B4
        a = [5 * B6, -6 * B5, 0, 0, 0, 0, ptol * b * b, -2 * ptol * b, ptol]
        abs_r = solve(a, 8, abs_r, 0.001)
        if rel0 is not None:
            a = [5 * B6, -6 * B5, 0, 0, rel_tol0 * b * b * m, -2 * rel_tol0 * b * m, rel_tol0 * m]
            rel0 = solve(a, 6, rel0, 0.001)
        if rel is not None:
            a = [5 * B6, -6 * B5, 0, 0, rel_tol * b * b * m, -2 * rel_tol * b * m, rel_tol * m]
            rel = solve(a, 6, rel, 0.001)
        outh = combine(abs_r, rel, rel0)
    return out0, outq, outh


def cell_norms(pos, mass):
    """(m, centre of mass, bmax, B2, B4, B6) of a set of bodies, FP64."""
    pos = np.asarray(pos, np.float64)
    mass = np.asarray(mass, np.float64)
    m = mass.sum()
    c = (mass[:, None] * pos).sum(0) / m
    r2 = ((pos - c) ** 2).sum(1)
    return (m, c, math.sqrt(r2.max()), float((mass * r2).sum()), float((mass * r2 ** 2).sum()),
            float((mass * r2 ** 3).sum()))


def bound(p, r, b, Bp1, Bp2):
    """Salmon-Warren bound E_p(r) on the acceleration error of the expansion
    through order p (p = 1, 2, 4 for the listed orders 0, 2, 4)."""
    return ((p + 2) * Bp1 / r ** (p + 1) - (p + 1) * Bp2 / r ** (p + 2)) / (r - b) ** 2


def accept_order(d, b, rc, count, quad_ncut=7, hexa_ncut=20, top=4):
    """Order (0, 2, 4) at which DLRcritMAC takes a cell at distance d from the
    sink (nln/mac.c:1784-1800), or None when it opens it; `top` caps the order."""
    if d > rc[0] + b:
        return 0
    if top >= 2 and quad_ncut and count >= quad_ncut and d > rc[1] + b:
        return 2
    if top >= 4 and hexa_ncut and count >= hexa_ncut and d > rc[2] + b:
        return 4
    return None
