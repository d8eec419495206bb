"""Pins the FP64 oracle (oracle/pgrav_oracle.c) before it is trusted as the checker.

Each oracle function is checked against (i) vectors the reference's own
compiled CPU kernels produced (tests/golden/ref_kernels.npz, and oracle/_ref
live when it is built), and (ii) what the mathematics fixes: convergence of the
multipole expansion to the direct sum at the rate (b/d)^(p+1), the single-body
special case, K1 continuity at r = eps, Newton's third law, and the Ewald
constants held in nln/ewald_le2.c:35-43.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _cell(rng, n=60, spread=1.0):
    pos = rng.normal(size=(n, 3)) * spread
    mass = rng.uniform(0.5, 1.5, n) / n
    cm = (pos * mass[:, None]).sum(0) / mass.sum()
    bmax = np.sqrt(((pos - cm) ** 2).sum(1)).max()
    return pos, mass, cm, bmax


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("kind,order,sz", [("H", 4, 29), ("Q", 2, 13), ("M", 0, 4)])
def test_oracle_matches_reference_vec_kernels(gold, kind, order, sz):
    """pHinteract/pQinteract/pMinteract (nln/pgrav_vec.c) vs orc_pmn: both use the
    physical (+) sign of the quadrupole potential."""
    recs, sinks = gold["cells"], gold["sinks"]
    f = oracle.to_soa(recs[:, :sz], 8)
    o = oracle.pmn(order, sinks, f, len(recs), sz, nsse=8, dipole=True)
    assert _rel(o, gold["vec_" + kind]) < 2e-6


def test_oracle_matches_reference_k1(gold):
    sinks, mono = gold["sinks"], gold["mono_k1"]
    o = oracle.pm_sk1(sinks, oracle.to_soa(mono, 8), mono.shape[0], 1.0 / gold["eps_k1"])
    assert _rel(o, gold["vec_MK1"]) < 2e-6


@pytest.mark.parametrize("name,order,sz,qsign", [("do_gravph", 4, 29, 1), ("do_gravpq", 2, 13, -1),
                                                 ("do_gravp", 0, 4, 1)])
def test_oracle_matches_reference_sse4(gold, name, order, sz, qsign):
    """Plummer-softened CPU kernels (nln/do_grav_sse4.c:1768-2224): no dipole,
    r^2 + eps^2, and do_gravpq's "phi -= eqe" (defect A2) via quad_phi_sign=-1."""
    recs, sinks = gold["cells"], gold["sinks"]
    src = recs[:, :sz] if sz != 4 else gold["mono_k1"]
    f4 = oracle.to_soa(src, 4)
    o = oracle.pmn(order, sinks, f4, src.shape[0], sz, nsse=4, dipole=False,
                   eps2=float(gold["eps2"]), quad_phi_sign=qsign)
    assert _rel(o, gold["sse4_" + name]) < 2e-6


def test_oracle_matches_reference_pgravm_sk1(gold):
    """_pgravm_sK1 (nln/py_pgrav/_pgrav.c) on the nln/py_pgrav/test_pgrav.py inputs."""
    src, sink = gold["tp_src"], gold["tp_sink"]
    seg = np.array([0, len(src)], np.uint32)
    o = oracle.pmm_sk1(sink[:, 1:4], seg, src[:, 1:4], float(src[0, 0]), 1.0 / 0.3)
    assert _rel(o, gold["tp_pgravm_sK1"]) < 1e-6


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
def test_oracle_vs_live_reference():
    rng = np.random.default_rng(7)
    recs = []
    for _ in range(16):
        pos, mass, cm, bmax = _cell(rng, 30, 0.1)
        pos += rng.uniform(-1, 1, 3)
        cm = (pos * mass[:, None]).sum(0) / mass.sum()
        recs.append(oracle.cell_record(pos, mass, cm, bmax))
    recs = np.array(recs)
    sinks = rng.uniform(-2, 2, (9, 3)).astype(np.float32)
    body = np.zeros((9, 14), np.float32)
    body[:, 1:4] = sinks
    f = oracle.to_soa(recs, 8)
    assert _rel(oracle.pmn(4, sinks, f, 16, 29), oracle.ref_mn("H", body, f, 2)) < 2e-6


# ------------------------------------------------------------ closed forms
def test_single_body_every_order_is_point_mass():
    rng = np.random.default_rng(1)
    p = rng.normal(size=(1, 3))
    rec = oracle.cell_record(p, np.array([0.7]), p[0], 1e-3)
    for order in (0, 2, 4):
        for dx in rng.normal(size=(5, 3)) * 3:
            r = np.linalg.norm(dx)
            a = oracle.mpole(rec, order, dx)
            np.testing.assert_allclose(a[3], -0.7 / r, rtol=1e-12)
            np.testing.assert_allclose(a[:3], -0.7 * dx / r ** 3, rtol=1e-12)


@pytest.mark.parametrize("order", [0, 2, 4])
def test_expansion_converges_at_order_rate(order):
    """|a_p - a_direct| ~ (b/d)^(p+2) relative to |a| ~ 1/d^2 (i.e. the truncation
    error of an order-p expansion falls as (b/d)^(p+1) relative).  A wrong sign,
    index or normalisation in any moment breaks the slope."""
    rng = np.random.default_rng(order + 11)
    pos, mass, cm, bmax = _cell(rng)
    rec = oracle.cell_record(pos, mass, cm, bmax)
    dirs = rng.normal(size=(24, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    errs = []
    for d in (6.0, 12.0, 24.0):
        e = []
        for u in dirs:
            x = cm + d * bmax * u
            ex = oracle.direct(x[None], pos, mass)[0]
            ap = oracle.mpole(rec, order, x - cm)
            e.append(np.linalg.norm(ap[:3] - ex[:3]) / np.linalg.norm(ex[:3]))
        errs.append(np.median(e))
    slopes = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    # first omitted term: order p+1, except that about the centre of mass the
    # dipole vanishes, so a monopole's error starts at the quadrupole
    expect = {0: 2, 2: 3, 4: 5}[order]
    assert np.all(np.abs(slopes - expect) < 0.45), (errs, slopes)


def test_potential_sign_of_quadrupole():
    """Defect A2: only the + sign converges to the direct-sum potential."""
    rng = np.random.default_rng(3)
    pos, mass, cm, bmax = _cell(rng)
    rec = oracle.cell_record(pos, mass, cm, bmax)
    x = cm + np.array([9.0, 4.0, -3.0]) * bmax
    ex = oracle.direct(x[None], pos, mass)[0]
    good = oracle.mpole(rec, 4, x - cm, quad_phi_sign=1)
    bad = oracle.mpole(rec, 4, x - cm, quad_phi_sign=-1)
    assert abs(good[3] - ex[3]) < 0.02 * abs(bad[3] - ex[3])
    np.testing.assert_allclose(good[:3], bad[:3], rtol=0, atol=0)


def test_records_are_traceless():
    rng = np.random.default_rng(5)
    pos, mass, cm, bmax = _cell(rng)
    q = oracle.cell_record(pos, mass, cm, bmax)
    # the stored components imply qzz = -(qxx+qyy) etc.; check against the full
    # traceless projection computed from raw moments
    u = oracle.raw_moments(pos, mass, cm)
    m2 = np.array([[u[4], u[5], u[6]], [u[5], u[7], u[8]], [u[6], u[8], u[9]]])
    t2 = (m2 - np.eye(3) * np.trace(m2) / 3) / bmax ** 2
    np.testing.assert_allclose([q[8], q[9], q[10], q[11], q[12]],
                               [t2[0, 0], t2[0, 1], t2[1, 1], t2[0, 2], t2[1, 2]], rtol=1e-12)
    # dipole about the centre of mass vanishes
    assert np.abs(q[5:8]).max() < 1e-14


def test_raw_moments_parallel_axis():
    """Second moments about c2 = moments about c1 shifted (parallel-axis identity,
    the relation nln/cofm.c:770-964 mpole_add applies)."""
    rng = np.random.default_rng(9)
    pos, mass, cm, _ = _cell(rng)
    c2 = cm + np.array([0.3, -0.2, 0.5])
    u1 = oracle.raw_moments(pos, mass, cm)
    u2 = oracle.raw_moments(pos, mass, c2)
    d = cm - c2
    M = u1[0]
    np.testing.assert_allclose(u2[1:4], u1[1:4] + M * d, atol=1e-13)
    np.testing.assert_allclose(u2[4], u1[4] + 2 * u1[1] * d[0] + M * d[0] ** 2, rtol=1e-12)
    np.testing.assert_allclose(u2[5], u1[5] + u1[1] * d[1] + u1[2] * d[0] + M * d[0] * d[1],
                               rtol=1e-12)


def test_k1_continuity_at_eps():
    """nln/pgrav.cu:760-766: u = 0 gives phi = -m/eps and a = -m x/eps^3."""
    eps, m = 0.25, 0.7
    for r in (eps * (1 - 1e-9), eps * (1 + 1e-9)):
        x = np.array([r, 0, 0], np.float32)
        o = oracle.pmm_sk1(np.zeros((1, 3), np.float32), np.array([0, 1], np.uint32),
                           -x[None], m, 1.0 / eps)[0]
        np.testing.assert_allclose(o[3], -m / eps, rtol=1e-6)
        np.testing.assert_allclose(o[0], -m * r / eps ** 3, rtol=1e-6)
    # inside: potential is finite at r = 0 and the force vanishes
    o = oracle.pmm_sk1(np.zeros((1, 3), np.float32), np.array([0, 1], np.uint32),
                       np.zeros((1, 3), np.float32), m, 1.0 / eps)[0]
    np.testing.assert_allclose(o[3], m / eps * (-45 / 32 - 3 / 8 - 1 / 2 - 1), rtol=1e-12)
    assert np.all(o[:3] == 0)


def test_direct_sum_two_body_and_third_law():
    rng = np.random.default_rng(2)
    p = rng.normal(size=(2, 3))
    mass = np.array([0.3, 0.9])
    o = oracle.direct(p, p, mass)
    d = p[1] - p[0]
    r = np.linalg.norm(d)
    np.testing.assert_allclose(o[0, :3], 0.9 * d / r ** 3, rtol=1e-13)
    np.testing.assert_allclose(o[0, 3], -0.9 / r, rtol=1e-13)
    np.testing.assert_allclose(mass[0] * o[0, :3], -mass[1] * o[1, :3], rtol=1e-13)
    ps = oracle.direct(p, p, mass, soft_type=1, eps=0.1)
    np.testing.assert_allclose(ps[0, 3], -0.9 / np.sqrt(r * r + 0.01) - 0.3 / 0.1, rtol=1e-13)


def test_ewald_constants_from_reference():
    """nln/ewald_le2.c:35-43: phi_ewald(x) + 1/|x| -> 2 x 1.4186487533 and
    f_ewald - x/|x|^3 -> -(4 pi / 3) x near the origin (unit box)."""
    x = np.array([1e-3, 4e-4, -3e-4])
    r = np.linalg.norm(x)
    f, phi = oracle.ewald(x)
    np.testing.assert_allclose(phi + 1 / r, 2 * 1.4186487533 - 2 * np.pi / 3 * r * r, rtol=2e-8)
    np.testing.assert_allclose((f - x / r ** 3) / x, -4.188790204786, rtol=2e-4)
    # symmetry: no force at the face and cube centres, odd in x
    for c in ([0.5, 0, 0], [0.5, 0.5, 0.5], [0, 0.5, 0.5]):
        assert np.abs(oracle.ewald(np.array(c))[0]).max() < 1e-12
    y = np.array([0.21, -0.13, 0.37])
    f1, p1 = oracle.ewald(y)
    f2, p2 = oracle.ewald(-y)
    np.testing.assert_allclose(f1, -f2, rtol=1e-12)
    np.testing.assert_allclose(p1, p2, rtol=1e-12)


def test_ewald_corr_exact_form_pins():
    """oracle.ewald_corr (nodes = 0), the checker of the periodic tree path:
    * values == the lattice sum of nln/ewald.c (oracle.ewald) minus the
      Newtonian term, at random points;
    * the closed-form Jacobian == central differences of those values, and
      dpsi/dr == C (psi is C's potential); the closed-form second derivatives
      == central differences of the Jacobian;
    * div C = -4 pi / L^3 and the Laplacian of C = 0 away from the lattice points
      (psi's Laplacian is the uniform background's);
    * at r = 0: psi(0) == 2 x 1.4186487533 (nln/ewald_le2.c:38, 43; the
      reference's literal is the simple cubic constant 2.837297479 to 1e-8),
      dC/dr == -(4 pi / 3) 1 (nln/ewald_le2.c:35-37), C == 0 and its second
      derivatives == 0;
    * odd C and even psi, and the box scaling C(L r) = C(r)/L^2."""
    rng = np.random.default_rng(5)
    r = rng.uniform(-0.5, 0.5, (64, 3))
    val, jac, third = oracle.ewald_corr(r)
    for i in range(16):
        f, phi = oracle.ewald(r[i])
        d = np.linalg.norm(r[i])
        np.testing.assert_allclose(val[i, :3], f - r[i] / d ** 3, rtol=0, atol=1e-13)
        np.testing.assert_allclose(val[i, 3], phi + 1 / d, rtol=0, atol=1e-13)
    h = 1e-5
    T = oracle.sym3_full(third)
    for b in range(3):
        e = np.zeros(3)
        e[b] = h
        vp, jp, _ = oracle.ewald_corr(r + e)
        vm, jm, _ = oracle.ewald_corr(r - e)
        fd = (vp - vm) / (2 * h)
        np.testing.assert_allclose(jac[:, :, b], fd, rtol=0, atol=1e-8 * np.abs(jac).max())
        np.testing.assert_allclose(T[:, :, :, b], (jp - jm)[:, :3, :] / (2 * h), rtol=0,
                                   atol=1e-8 * np.abs(T).max())
    np.testing.assert_allclose(jac[:, 3, :], val[:, :3], rtol=0, atol=1e-13)
    np.testing.assert_allclose(np.trace(jac[:, :3, :], axis1=1, axis2=2), -4 * np.pi, rtol=1e-10)
    np.testing.assert_allclose(np.einsum("nijj->ni", T), 0, atol=1e-9 * np.abs(T).max())
    v0, j0, t0 = oracle.ewald_corr(np.zeros((1, 3)))
    np.testing.assert_allclose(v0[0, 3], 2 * 1.4186487533, rtol=2e-8)
    np.testing.assert_allclose(v0[0, 3], 2.837297479, rtol=1e-9)
    np.testing.assert_allclose(j0[0, :3, :], -4 * np.pi / 3 * np.eye(3), rtol=0, atol=1e-12)
    assert np.abs(v0[0, :3]).max() < 1e-15 and np.abs(t0).max() < 1e-12
    vn, jn, tn = oracle.ewald_corr(-r)
    np.testing.assert_allclose(vn[:, :3], -val[:, :3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(vn[:, 3], val[:, 3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(tn, -third, rtol=0, atol=1e-10)
    v2, _, _ = oracle.ewald_corr(2.0 * r[:8], L=2.0)
    np.testing.assert_allclose(v2[:, :3], val[:8, :3] / 4, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v2[:, 3], val[:8, 3] / 2, rtol=0, atol=1e-12)


def test_ewald_corr_table_form_pins():
    """The tabulated form (nearest node, Taylor steps with the node's value,
    Jacobian and second derivatives): exact at the nodes; its error against the
    exact sum falls faster than h^2 when the node spacing halves and, at the 64
    nodes per half box the GPU tables hold, stays below 5e-6 of max |C| over the
    box; the step from a node is the Taylor polynomial of its derivatives."""
    rng = np.random.default_rng(6)
    r = rng.uniform(-0.5, 0.5, (400, 3))
    exact, _, _ = oracle.ewald_corr(r)
    errs = {}
    for nodes in (17, 33, 65):
        tab, _, _ = oracle.ewald_corr(r, nodes=nodes)
        errs[nodes] = np.abs(tab - exact).max()
    assert errs[17] / errs[33] > 4.0 and errs[33] / errs[65] > 4.0, errs
    tab64, j64, t64 = oracle.ewald_corr(r, nodes=64)
    assert np.abs(tab64 - exact).max() < 5e-6 * np.abs(exact[:, :3]).max()
    h = 0.5 / 63
    nodes = np.rint(r / h) * h
    vn, jn, tn = oracle.ewald_corr(nodes)
    tv, tj, tt = oracle.ewald_corr(nodes, nodes=64)
    np.testing.assert_allclose(tv, vn, rtol=0, atol=1e-14)
    np.testing.assert_allclose(tt, tn, rtol=0, atol=0)
    np.testing.assert_array_equal(t64, tn)  # the node's second derivatives
    d = r - nodes
    T = oracle.sym3_full(tn)
    Td = np.einsum("nabc,nc->nab", T, d)
    Tdd = np.einsum("nab,nb->na", Td, d)
    c = vn[:, :3] + np.einsum("nab,nb->na", jn[:, :3, :], d) + 0.5 * Tdd
    psi = (vn[:, 3] + (vn[:, :3] * d).sum(1) + 0.5 * np.einsum("na,nab,nb->n", d, jn[:, :3, :], d)
           + (Tdd * d).sum(1) / 6)
    np.testing.assert_allclose(tab64[:, :3], c, rtol=0, atol=1e-13)
    np.testing.assert_allclose(tab64[:, 3], psi, rtol=0, atol=1e-13)
    np.testing.assert_allclose(j64[:, :3, :], jn[:, :3, :] + Td, rtol=0, atol=1e-12)


@pytest.mark.parametrize("soft,eps", [(0, 0.0), (1, 0.05)])
def test_direct_grad_matches_finite_differences_of_direct(soft, eps):
    """oracle.direct_grad (the adjoint's FP64 checker) against central
    differences of oracle.direct: L = sum_i w_i . a_i + v_i phi_i over every
    body, each coordinate and mass nudged by h = 1e-6 (relative), unsoftened and
    Plummer-softened."""
    rng = np.random.default_rng(11 + soft)
    n = 24
    pos = rng.normal(size=(n, 3))
    mass = rng.random(n) + 0.2
    w4 = rng.normal(size=(n, 4))
    g = oracle.direct_grad(pos, mass, w4, soft_type=soft, eps=eps)

    def loss(p, m):
        r = oracle.direct(p, p, m, soft_type=soft, eps=eps)
        return (w4 * r).sum()

    h = 1e-6
    fd = np.zeros((n, 4))
    for i in range(n):
        for k in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, k] += h
            pm[i, k] -= h
            fd[i, k] = (loss(pp, mass) - loss(pm, mass)) / (2 * h)
        mp, mm = mass.copy(), mass.copy()
        mp[i] += h
        mm[i] -= h
        fd[i, 3] = (loss(pos, mp) - loss(pos, mm)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * np.abs(fd).max())
