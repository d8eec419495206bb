"""TEST INFRASTRUCTURE ONLY -- numpy restatement of the reference's periodic
far field for explicit images: the cubic local expansion of the images beyond
the (2 nimage + 1)^3 the walk lists (nln/ewald_le2.c: o4 :49-85, o6 :87-183,
o8 :185-337, ewald_le :339-381) and the uniform-background terms
(ewald_background :26-48), with the scales of FixCubeEwaldLE
(nln/main.c:1811-1862).  Only tests/ import it, as the checker.

The reference evaluates, per body at x (box units, box centred on 0),

    phi += t_k * sum_j (m_j / M) K_k(x - x_j)        k = 4, 6, 8

through hand-expanded moment polynomials (the Q[] of :688-771), where K_k is
the cubic-invariant harmonic of degree k and t_k a combination of eight
lattice sums hard-coded for nimage = 1 and 2 (:347-365, the coefficient lines
:65, :111, :217).  This restatement evaluates the same sum directly over body
pairs, O(N^2), from the constants and the three polynomials; it shares no
code with the library's moment expansion (host/ewald_le.cpp), which computes
the lattice sums itself.

parity pins (tests/test_ewald_le_cpu.py): the constants are the reference's
(golden fixture tests/golden/ewald_le2_constants.json); the polynomials are
checked harmonic and cubic-invariant and against the coefficients of the
reference's Q[1], Q[14], Q[31]; the whole scheme (near images summed directly
+ this far field) against the Ewald sum restated from nln/ewald.c.
"""
import itertools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden", "ewald_le2_constants.json")

# nln/ewald_le2.c:38-43: 2 x 1.4186487533, the lattice self constant
XI_REF = 2.0 * 1.4186487533


def constants():
    with open(GOLDEN) as f:
        return json.load(f)


def lattice_t(nimage):
    """(t4, t6, t8) of nln/ewald_le2.c:65, :111, :217 from the reference's constants."""
    e = constants()["ewx"][str(nimage)]
    t4 = 5.25 * e["ewx2y2"] - 1.75 * e["ewx4"]
    t6 = -11.25 * e["ewx2y2z2"] + 5.625 * e["ewx4y2"] - 0.375 * e["ewx6"]
    t8 = -54.140625 * e["ewx4y4"] + 43.3125 * e["ewx6y2"] - 1.546875 * e["ewx8"]
    return t4, t6, t8


def _sym(terms):
    """Monomial dict from {sorted exponent tuple: coefficient}, all axis permutations."""
    out = {}
    for e, c in terms.items():
        for p in set(itertools.permutations(e)):
            out[p] = c
    return out


# cubic-invariant harmonic polynomials, unit x^k coefficient
K = {4: _sym({(4, 0, 0): 1.0, (2, 2, 0): -3.0}),
     6: _sym({(6, 0, 0): 1.0, (4, 2, 0): -7.5, (2, 2, 2): 90.0}),
     8: _sym({(8, 0, 0): 1.0, (6, 2, 0): -14.0, (4, 4, 0): 35.0})}


def poly(terms, r):
    """Value and gradient of sum c x^a y^b z^c at r (..., 3)."""
    x, y, z = r[..., 0], r[..., 1], r[..., 2]
    v = np.zeros(x.shape)
    g = np.zeros(r.shape)
    for (a, b, c), k in terms.items():
        v += k * x ** a * y ** b * z ** c
        if a:
            g[..., 0] += k * a * x ** (a - 1) * y ** b * z ** c
        if b:
            g[..., 1] += k * b * x ** a * y ** (b - 1) * z ** c
        if c:
            g[..., 2] += k * c * x ** a * y ** b * z ** (c - 1)
    return v, g


def far_field(pos, mass, box=1.0, nimage=2, G=1.0):
    """(n, 4) = (a, phi) that FixCubeEwaldLE adds to each body (nln/main.c:1836-1860)
    for bodies in [0, box)^3: ewald_background (not the subtract_background mode)
    plus ewald_le, in FP64."""
    pos = np.asarray(pos, np.float64)
    mass = np.asarray(mass, np.float64)
    M = mass.sum()
    x = (pos - 0.5 * box) / box              # nln/main.c:1841-1842, 1852-1853
    t = dict(zip((4, 6, 8), lattice_t(nimage)))
    out = np.zeros((pos.shape[0], 4))
    for i in range(pos.shape[0]):
        d = x[i] - x                         # x - x_j
        pp = 0.0
        grad = np.zeros(3)
        for k in (8, 6, 4):                  # nln/ewald_le2.c:365-367
            v, g = poly(K[k], d)
            pp += t[k] * (mass * v).sum() / M
            grad += t[k] * (mass[:, None] * g).sum(0) / M
        # ewald_background (nln/ewald_le2.c:26-48); phicorr of nln/main.c:1826-1834
        q = (mass[:, None] * x).sum(0) / M
        q2 = (mass * (x ** 2).sum(1)).sum() / M
        phicorr = sum(1.0 / np.sqrt(a * a + b * b + c * c)
                      for a in range(-nimage, nimage + 1) for b in range(-nimage, nimage + 1)
                      for c in range(-nimage, nimage + 1) if a or b or c)
        bg_acc = 4.0 * np.pi / 3.0 * (x[i] - q)
        bg_phi = (phicorr + XI_REF - 2.0 * np.pi / 3.0 * (x[i] @ x[i] - 2 * x[i] @ q + q2)
                  - XI_REF * mass[i] / M)
        gm = G * M
        out[i, :3] = gm / box ** 2 * (bg_acc - grad)  # a -= t dP/dx (ewald_le2.c:68-84)
        out[i, 3] = gm / box * (bg_phi + pp)
    return out


def far_field_sb(pos, mass, box=1.0, nimage=2, G=1.0):
    """The subtract_background form (FixCubeEwaldLE with subtract_background,
    nln/main.c:1839-1860): the sources are the bodies less a uniform cube of their
    total mass filling the box (calculate_cartesian_moments with msb adds the unit
    cube's moments times -M, nln/ewald_le2.c:605-646, and sets Q[0] = 0, :688-691),
    no ewald_background, and the acceleration less (4 pi / 3) times the normalised
    dipole (nln/ewald_le2.c:371-375).  Here the cube's part of each sum is its
    integral, by a 5-point Gauss-Legendre product rule (exact for degree 8)."""
    pos = np.asarray(pos, np.float64)
    mass = np.asarray(mass, np.float64)
    M = mass.sum()
    x = (pos - 0.5 * box) / box
    t = dict(zip((4, 6, 8), lattice_t(nimage)))
    gx, gw = np.polynomial.legendre.leggauss(5)
    gx, gw = 0.5 * gx, 0.5 * gw
    cube = np.array(list(itertools.product(gx, gx, gx)))
    cw = np.array([a * b * c for a, b, c in itertools.product(gw, gw, gw)])
    q = (mass[:, None] * x).sum(0) / M
    out = np.zeros((pos.shape[0], 4))
    for i in range(pos.shape[0]):
        pp = 0.0
        grad = np.zeros(3)
        for k in (8, 6, 4):
            v, g = poly(K[k], x[i] - x)
            vc, gc = poly(K[k], x[i] - cube)
            pp += t[k] * ((mass * v).sum() / M - (cw * vc).sum())
            grad += t[k] * ((mass[:, None] * g).sum(0) / M - (cw[:, None] * gc).sum(0))
        gm = G * M
        out[i, :3] = gm / box ** 2 * (-grad - 4.0 * np.pi / 3.0 * q)
        out[i, 3] = gm / box * pp
    return out


def near_images(pos, mass, box=1.0, nimage=2, G=1.0):
    """FP64 direct sum over the (2 nimage + 1)^3 explicit images of every body (the
    walk's coverage, nln/mac.c:91-111, 734), the self pair of the central image
    excluded; phi = -G m / r (nln/grav_n2.c)."""
    pos = np.asarray(pos, np.float64)
    mass = np.asarray(mass, np.float64)
    rng = range(-nimage, nimage + 1)
    out = np.zeros((pos.shape[0], 4))
    for off in itertools.product(rng, rng, rng):
        src = pos + box * np.asarray(off, np.float64)
        d = src[None, :, :] - pos[:, None, :]
        r2 = (d ** 2).sum(-1)
        if off == (0, 0, 0):
            np.fill_diagonal(r2, np.inf)
        rinv = 1.0 / np.sqrt(r2)
        out[:, :3] += G * (mass[None, :, None] * d * (rinv ** 3)[..., None]).sum(1)
        out[:, 3] -= G * (mass[None, :] * rinv).sum(1)
    return out
