"""The error-tolerance acceptance test (mac = 1): the reference's AREL_MAC
critical radii (nln/cofm.c:276-565) and the DLRcritMAC order choice
(nln/mac.c:1776-1800), restated in oracle/mac_arel.py.

Pins, beyond the host == oracle comparison of the two restatements:
- the radii bound the error they promise: at r = rcrit of each order, the
  actual acceleration error of that order's expansion of the cell (evaluated by
  the oracle's own multipole code against the exact sum over the cell's bodies)
  is below the tolerance in every sampled direction, and not far below it in the
  worst case (a bound 10x too loose fails);
- Newton's root: the Salmon-Warren bound at rcrit is within the margin below
  the tolerance;
- the lists: every listed cell is cleared by every sink of the group at its
  order's radius + bmax, lower orders are not, and each sink's total error
  stays under (cells listed) x errtol against the FP64 direct sum.
"""
import math

import numpy as np
import pytest

import listoracle
import oracle
from oracle import mac_arel as MA
from pgrav import batch, ics

ORDERS = (0, 2, 4)
P_OF = {0: 1, 2: 2, 4: 4}  # expansion through order p: Salmon-Warren index


def _cells(rng, count):
    for trial in range(count):
        n = int(rng.integers(8, 300))
        if trial % 2:
            pos = rng.uniform(-1, 1, (n, 3)) * rng.uniform(0.2, 1, 3)
        else:
            pos = rng.normal(size=(n, 3)) * rng.uniform(0.1, 1, 3)
        yield pos, rng.uniform(0.5, 1.5, n)


def _true_error(pos, mass, order, r, ndir=64, seed=0):
    m, c, b, *_ = MA.cell_norms(pos, mass)
    rec = oracle.cell_record(pos, mass, c, b if b > 0 else 1.0)
    u = np.random.default_rng(seed).normal(size=(ndir, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    sinks = c + r * u
    approx = oracle.eval_records(sinks, rec[None, :], order, dipole=False)
    dx = sinks[:, None, :] - pos[None, :, :]
    d = np.linalg.norm(dx, axis=2)
    exact = -(mass[None, :, None] * dx / d[..., None] ** 3).sum(1)
    return float(np.linalg.norm(approx[:, :3] - exact, axis=1).max())


@pytest.mark.parametrize("frac_tol,frac_tol0,ptol_boost,stol_max",
                         [(0.0, 0.0, 0.0, 0.0), (0.02, 0.0, 0.0, 0.0), (0.05, 1e-4, 0.0, 0.0),
                          (0.0, 0.0, 3.0, 0.0), (0.0, 0.0, 0.0, 50.0)])
def test_host_radii_equal_oracle(frac_tol, frac_tol0, ptol_boost, stol_max):
    rng = np.random.default_rng(1)
    for pos, mass in _cells(rng, 40):
        mass = mass * 10 ** rng.uniform(0, 12)  # spans the sigma(m) table
        m, c, b, B2, B4, B6 = MA.cell_norms(pos, mass)
        tol = 10 ** rng.uniform(-6, -1) * m / b ** 2
        kw = dict(errtol=tol, frac_tol=frac_tol, frac_tol0=frac_tol0, ptol_boost=ptol_boost,
                  stol_max=stol_max, quad_ncut=7, hexa_ncut=20)
        want = MA.rcrit(m, b, B2, B4, B6, len(mass), tol, frac_tol, frac_tol0, ptol_boost, 2.0,
                        stol_max, 7, 20)
        got = batch.arel_rcrit(m, b, B2, B4, B6, len(mass), 2.0, **kw)
        for g, w in zip(got, want):
            if math.isnan(w):
                assert math.isnan(g)
            else:
                assert abs(g - w) <= 1e-12 * w, (got, want)


def test_radii_bound_the_true_error():
    rng = np.random.default_rng(2)
    worst = {o: 0.0 for o in ORDERS}
    for pos, mass in _cells(rng, 60):
        m, c, b, B2, B4, B6 = MA.cell_norms(pos, mass)
        tol = 10 ** rng.uniform(-5, -1) * m / b ** 2
        rc = MA.rcrit(m, b, B2, B4, B6, len(mass), tol, quad_ncut=1, hexa_ncut=1)
        for k, order in enumerate(ORDERS):
            assert rc[k] > b
            ratio = _true_error(pos, mass, order, rc[k]) / tol
            assert ratio <= 1.0, (order, ratio)
            worst[order] = max(worst[order], ratio)
    # the bound is not loose by 10x for every cell
    assert min(worst.values()) > 0.1, worst


def test_radii_are_roots_of_the_bound():
    rng = np.random.default_rng(3)
    for pos, mass in _cells(rng, 30):
        m, c, b, B2, B4, B6 = MA.cell_norms(pos, mass)
        tol = 10 ** rng.uniform(-5, -1) * m / b ** 2
        rc = MA.rcrit(m, b, B2, B4, B6, len(mass), tol, quad_ncut=1, hexa_ncut=1)
        # order 0 uses the lower bound B2^1.5 / m^0.5 for B3; orders 2 and 4 bound
        # B3 and B5 by bmax B2 and bmax B4
        B3 = B2 * math.sqrt(B2 / m)
        terms = {0: (B2, B3), 2: (b * B2, B4), 4: (b * B4, B6)}
        for k, order in enumerate(ORDERS):
            e = MA.bound(P_OF[order], rc[k], b, *terms[order])
            assert 0.5 * tol < e <= tol, (order, e / tol)


def test_single_body_and_coincident_cells_have_zero_radius():
    got = batch.arel_rcrit(3.0, 0.0, 0.0, 0.0, 0.0, 1, 1.0, errtol=1e-3)
    assert got[0] == 0.0 and math.isnan(got[1]) and math.isnan(got[2])
    got = batch.arel_rcrit(30.0, 0.0, 0.0, 0.0, 0.0, 30, 1.0, errtol=1e-3)
    assert list(got) == [0.0, 0.0, 0.0]


def test_bad_tolerances_raise():
    pos, mass = ics.plummer(500, seed=3)
    for kw in (dict(errtol=0.0), dict(errtol=1e-3, frac_tol=-1.0), dict(errtol=1e-3, hexa_ncut=-2)):
        with pytest.raises(batch.PgravError):
            batch.HostPlan(pos, mass, mac="tolerance", **kw)
    with pytest.raises(ValueError):
        batch.make_params(mac="bogus")
    # the other decomposition parameters are checked on the same path
    for kw in (dict(merge_groups=6), dict(dispatch=4), dict(leaf_max=-1), dict(mixed_order=2),
               dict(group_box=3), dict(group_max=65), dict(order=3)):
        with pytest.raises(batch.PgravError):
            batch.HostPlan(pos, mass, **kw)


@pytest.fixture(scope="module")
def arel_plan():
    pos, mass = ics.clustered(12000, seed=21, nhalo=12)
    errtol = 0.1  # about 1e-3 of the median acceleration of this set (~1e2)
    kw = dict(order=4, eps=1e-3, soft="plummer", leaf_max=32, mac="tolerance", errtol=errtol,
              frac_tol=0.02)
    return pos, mass, kw, batch.HostPlan(pos, mass, **kw)


def test_lists_follow_the_acceptance_order(arel_plan):
    pos, mass, kw, H = arel_plan
    rc = H.rcrit()
    perm, spos = H.sorted()
    st = H.stats()
    assert st["pairs_mono"] > 0 and st["pairs_quad"] > 0 and st["pairs_hexa"] > 0
    _, cnt, cen = H.groups()
    rng = np.random.default_rng(4)
    kinds = {"mono": 0, "quad": 2, "hexa": 4}
    for g in rng.choice(st["ngroups"], 40, replace=False):
        sinks = spos[H.group_slots(int(g))]
        for kind, order in kinds.items():
            for nd in H.list(int(g), kind):
                c = H.cell(int(nd))
                b = c["bmax"]
                k = ORDERS.index(order)
                d_all = np.linalg.norm(sinks - c["cofm"], axis=1)
                # every sink clears this order's radius
                assert d_all.min() > rc[nd, k] + b
                # the group's nearest possible sink does not clear a lower order
                D = np.linalg.norm(cen[g, :3] - c["cofm"]) - cen[g, 3]
                want = MA.accept_order(D, b, rc[nd], c["count"], 7, 20, top=4)
                assert want == order, (g, nd, want, order)


def test_errors_stay_under_the_tolerance_sum(arel_plan):
    pos, mass, kw, H = arel_plan
    perm, spos = H.sorted()
    smass = mass[perm]
    st = H.stats()
    rng = np.random.default_rng(5)
    for g in rng.choice(st["ngroups"], 8, replace=False):
        _, got = listoracle.group_reference(H, int(g), spos, smass, soft="plummer", eps=1e-3)
        sinks = spos[H.group_slots(int(g))]
        exact = oracle.direct(sinks, spos, smass, soft_type=1, eps=1e-3)
        ncell = sum(H.list(int(g), k).size for k in ("mono", "quad", "hexa"))
        err = np.linalg.norm(got[:, :3] - exact[:, :3], axis=1)
        assert err.max() <= ncell * kw["errtol"], (g, err.max(), ncell)


def test_tighter_tolerance_lists_more():
    pos, mass = ics.clustered(12000, seed=21, nhalo=12)
    sizes = []
    for et in (1.0, 0.1, 0.01):
        H = batch.HostPlan(pos, mass, order=4, eps=1e-3, soft="plummer", leaf_max=32,
                           mac="tolerance", errtol=et)
        sizes.append(H.stats()["pairs_total"])
    assert sizes[0] < sizes[1] < sizes[2], sizes


@pytest.mark.parametrize("group_box", [0, 1])
def test_opening_angle_holds_for_every_sink(group_box):
    """theta MAC, measured from the group's enclosing sphere (0) or its bounding
    box (1): every sink of a group is farther than bmax / theta from each listed
    cell's centre of mass; the box lists fewer pairs."""
    pos, mass = ics.clustered(12000, seed=21, nhalo=12)
    kw = dict(theta=0.6, order=4, eps=1e-3, soft="plummer", merge_groups=5, rcb_cell=1024,
              leaf_max=32)
    H = batch.HostPlan(pos, mass, group_box=group_box, **kw)
    perm, spos = H.sorted()
    st = H.stats()
    rng = np.random.default_rng(6)
    for g in rng.choice(st["ngroups"], 40, replace=False):
        sinks = spos[H.group_slots(int(g))]
        for nd in H.list(int(g), "hexa"):
            c = H.cell(int(nd))
            d = np.linalg.norm(sinks - c["cofm"], axis=1)
            assert d.min() > c["bmax"] / 0.6
    if group_box:
        sphere = batch.HostPlan(pos, mass, group_box=0, **kw).stats()["pairs_total"]
        assert st["pairs_total"] < sphere
