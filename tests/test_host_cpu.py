"""Host half of the batched path (tree, FP64 moments, groups, lists, ICs) on
the CPU: the invariants the reference's walk relies on, checked against the
oracle where the oracle has the answer."""
import numpy as np
import pytest

import listoracle
import oracle
from pgrav import batch, ics
from pgrav._lib import PgravError


@pytest.fixture(scope="module")
def clustered():
    pos, mass = ics.clustered(30000, seed=11, nhalo=32)
    H = batch.HostPlan(pos, mass, theta=0.6, order=4, eps=1e-4)
    perm, spos = H.sorted()
    return pos, mass, H, perm, spos


def test_sort_is_a_permutation(clustered):
    pos, mass, H, perm, spos = clustered
    assert np.array_equal(np.sort(perm), np.arange(pos.shape[0]))
    np.testing.assert_array_equal(spos, pos[perm])


def test_moment_translation_matches_direct_moments(clustered):
    """FP64 M2M translation (nln/cofm.c:770-964) == moments summed directly
    from the cell's bodies about its centre of mass (parallel-axis identity)."""
    pos, mass, H, perm, spos = clustered
    smass = mass[perm]
    st = H.stats()
    rng = np.random.default_rng(0)
    for nd in [0, 1, 2] + list(rng.integers(0, st["nnodes"], 60)):
        c = H.cell(int(nd))
        b = spos[c["first"]:c["first"] + c["count"]]
        m = smass[c["first"]:c["first"] + c["count"]]
        cm = (b * m[:, None]).sum(0) / m.sum()
        np.testing.assert_allclose(c["cofm"], cm, rtol=0, atol=1e-13)
        raw = oracle.raw_moments(b, m, c["cofm"])
        scale = np.abs(raw).max()
        assert np.abs(raw - c["raw"]).max() <= 1e-10 * scale
        true_bmax = np.sqrt(((b - cm) ** 2).sum(1)).max()
        assert c["bmax"] >= true_bmax * (1 - 1e-12) - 1e-14  # 1-body cells: 0 vs round-off


def test_groups_partition_bodies(clustered):
    pos, mass, H, perm, spos = clustered
    st = H.stats()
    nxt = 0
    for g in range(st["ngroups"]):
        first, cnt, c = H.group(g)
        assert first == nxt and 1 <= cnt <= 64
        b = spos[first:first + cnt]
        assert np.sqrt(((b - c[:3]) ** 2).sum(1)).max() <= c[3] * (1 + 1e-12)
        nxt += cnt
    assert nxt == pos.shape[0]


def test_lists_cover_every_body_once_and_obey_the_mac(clustered):
    pos, mass, H, perm, spos = clustered
    st = H.stats()
    rng = np.random.default_rng(1)
    for g in rng.choice(st["ngroups"], 25, replace=False):
        assert listoracle.interactions_cover_all(H, int(g), pos.shape[0])
        _, _, c = H.group(int(g))
        for nd in H.list(int(g), "hexa"):
            cell = H.cell(int(nd))
            d = np.linalg.norm(cell["cofm"] - c[:3])
            assert d - c[3] > cell["bmax"] / 0.6
        body = H.list(int(g), "body")
        assert np.unique(body).size == body.size


def test_build_is_deterministic():
    pos, mass = ics.clustered(8000, seed=5, nhalo=8)
    a = batch.HostPlan(pos, mass, theta=0.5, order=4, eps=1e-3)
    b = batch.HostPlan(pos, mass, theta=0.5, order=4, eps=1e-3, nthreads=3)
    assert a.stats()["pairs_total"] == b.stats()["pairs_total"]
    for g in range(0, a.stats()["ngroups"], 7):
        for kind in batch.KINDS:
            np.testing.assert_array_equal(a.list(g, kind), b.list(g, kind))
    np.testing.assert_array_equal(a.lpt(), b.lpt())


@pytest.mark.parametrize("order,kind", [(2, "quad"), (0, "mono")])
def test_lower_orders_fill_their_lists(order, kind):
    pos, mass = ics.plummer(6000, seed=3)
    H = batch.HostPlan(pos, mass, theta=0.6, order=order, eps=0.01)
    st = H.stats()
    assert st["entries_" + kind] > 0 and st["entries_hexa"] == 0


def test_group_modes():
    pos, mass = ics.clustered(20000, seed=6, nhalo=16)
    chunks = batch.HostPlan(pos, mass, merge_groups=2).stats()
    assert chunks["ngroups"] == (20000 + 63) // 64
    merged = batch.HostPlan(pos, mass, merge_groups=1).stats()
    cells = batch.HostPlan(pos, mass, merge_groups=0).stats()
    assert merged["mean_group"] > cells["mean_group"]


def test_periodic_lists_use_nearest_images():
    pos, mass = ics.lattice_grf(16, seed=2)
    H = batch.HostPlan(pos, mass, periodic=True, box=1.0, theta=0.6, order=4, eps=1e-3)
    st = H.stats()
    for g in range(0, st["ngroups"], 5):
        assert listoracle.interactions_cover_all(H, g, pos.shape[0])


def test_reference_layout_emitters(clustered):
    """SoA-8 blocks (nln/pgrav.cu:66 indexing) and MMvec segments built from a
    group's lists hold exactly the list's records and bodies."""
    from pgrav import reflists
    pos, mass, H, perm, spos = clustered
    g = H.stats()["ngroups"] // 2
    soa, n = reflists.group_soa(H, g, "hexa")
    nodes = H.list(g, "hexa")
    assert n % 8 == 0 and n >= nodes.size
    recs = H.records(nodes)
    i = nodes.size - 1
    for k in (0, 1, 4, 12, 28):
        assert soa[(i // 8) * 8 * 29 + k * 8 + i % 8] == np.float32(recs[i, k])
    seg, nb = reflists.group_segments(H, g)
    body = np.sort(H.list(g, "body"))
    assert nb == body.size
    expand = np.concatenate([np.arange(b, b + l) for b, l in seg])
    np.testing.assert_array_equal(expand, body)


def test_bad_parameters_raise():
    pos, mass = ics.plummer(100, seed=1)
    with pytest.raises(PgravError):
        batch.HostPlan(pos, mass, group_max=65)
    with pytest.raises(PgravError):
        batch.HostPlan(pos, mass, order=3)


def test_ics_shapes_and_statistics():
    pos, mass = ics.plummer(200000, seed=9)
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 10.0 + 1e-9
    # truncated-Plummer median: X = 1/2 of the retained mass fraction
    xmax = 1.01 ** -1.5
    r_half = 1.0 / np.sqrt((0.5 * xmax) ** (-2 / 3) - 1)
    assert abs(np.median(r) / r_half - 1) < 0.02
    np.testing.assert_allclose(mass.sum(), 1.0)
    pos, mass = ics.clustered(50000, seed=1, nhalo=64)
    inside = np.all((pos >= 0) & (pos < 1), axis=1).mean()
    assert 0.2 <= inside <= 1.0
    pos, mass = ics.lattice_grf(32, seed=3, rms=0.2)
    assert pos.shape == (32 ** 3, 3) and np.all((pos >= 0) & (pos < 1))
    g = (np.arange(32) + 0.5) / 32
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    d = pos - np.stack([gx, gy, gz], -1).reshape(-1, 3)
    d -= np.rint(d)
    np.testing.assert_allclose(np.sqrt((d ** 2).sum(1).mean()) * 32, 0.2, rtol=1e-3)


def test_mixed_order_lists_keep_the_truncation_bound():
    """mixed_order: the same accepted cells, each listed at the lowest order whose
    (bmax/d')^(q+1) bound (d' = d - r_group) is within theta^5; evaluated by the
    FP64 list oracle the force error stays at the single-order level."""
    pos, mass = ics.clustered(12000, seed=21, nhalo=12)
    theta = 0.6
    one = batch.HostPlan(pos, mass, theta=theta, order=4, eps=1e-3)
    mix = batch.HostPlan(pos, mass, theta=theta, order=4, eps=1e-3, mixed_order=True)
    s1, s2 = one.stats(), mix.stats()
    # identical acceptance: only the order of accepted cells changes
    assert s1["ngroups"] == s2["ngroups"] and s1["entries_body"] == s2["entries_body"]
    assert s2["entries_hexa"] + s2["entries_quad"] + s2["entries_mono"] == s1["entries_hexa"]
    assert s2["entries_quad"] > 0.3 * s1["entries_hexa"]
    perm, spos = mix.sorted()
    smass = mass[perm]
    rng = np.random.default_rng(4)
    e1, e2 = [], []
    for g in rng.choice(s2["ngroups"], 12, replace=False):
        g = int(g)
        _, _, c = mix.group(g)
        for kind, q in (("quad", 2), ("mono", 0)):
            for nd in mix.list(g, kind):
                cell = mix.cell(int(nd))
                d = np.linalg.norm(cell["cofm"] - c[:3]) - c[3]
                assert (cell["bmax"] / d) ** (q + 1) <= theta ** 5 * (1 + 1e-12)
        f, r1 = listoracle.group_reference(one, g, spos, smass, eps=1e-3)
        _, r2 = listoracle.group_reference(mix, g, spos, smass, eps=1e-3)
        ex = oracle.direct(spos[f:f + r1.shape[0]], spos, smass, soft_type=1, eps=1e-3)
        n = np.linalg.norm(ex[:, :3], axis=1)
        e1.append(np.linalg.norm(r1[:, :3] - ex[:, :3], axis=1) / n)
        e2.append(np.linalg.norm(r2[:, :3] - ex[:, :3], axis=1) / n)
    m1, m2 = np.median(np.concatenate(e1)), np.median(np.concatenate(e2))
    assert m2 < 3 * m1 and m2 < 1e-3, (m1, m2)


@pytest.mark.parametrize("rcb_cell", [512, 1 << 20])
def test_bisection_groups_partition_bodies_in_balanced_sets(rcb_cell):
    """merge_groups 5: tree cells of <= rcb_cell bodies cut into ceil(n/64)
    groups of near-equal size by coordinate bisection; the groups partition the
    bodies, every sink's lists still cover N once, and the group sphere holds
    every member."""
    pos, mass = ics.clustered(20000, seed=12, nhalo=16)
    H = batch.HostPlan(pos, mass, theta=0.6, order=4, eps=1e-4, merge_groups=5, leaf_max=32,
                       rcb_cell=rcb_cell)
    perm, spos = H.sorted()
    st = H.stats()
    seen = np.zeros(pos.shape[0], np.int64)
    sizes = []
    for g in range(st["ngroups"]):
        s = H.group_slots(g)
        seen[s] += 1
        sizes.append(s.size)
        _, cnt, c = H.group(g)
        assert cnt == s.size <= 64
        assert (np.linalg.norm(spos[s] - c[:3], axis=1) <= c[3] * (1 + 1e-12)).all()
    assert (seen == 1).all()
    sizes = np.array(sizes)
    if rcb_cell > pos.shape[0]:  # one cell: every group within one body of the others
        assert sizes.max() - sizes.min() <= 1 and sizes.min() >= 32
        assert sizes.mean() > batch.HostPlan(pos, mass, theta=0.6, order=4, eps=1e-4,
                                             merge_groups=4, leaf_max=32).stats()["mean_group"]
    for g in range(0, st["ngroups"], max(1, st["ngroups"] // 20)):
        assert listoracle.interactions_cover_all(H, g, pos.shape[0])
