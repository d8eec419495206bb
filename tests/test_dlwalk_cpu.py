"""The DL walk's lists for the drop-in path (pgrav_dl_*), on the CPU.

1. The per-call oracle (oracle/dlmac.py) against known answers worked by hand
   from nln/mac.c:1773-1844 on a two-cell tree (acceptance order M -> Q -> H,
   the dlfac split, waiting body sources, the terminal expansion and its gate).
2. Every MAC call the library's walk makes, replayed through the oracle: the
   same result per source and the same appends in the same order.
3. Every body accounts for every body exactly once (nimage N bodies; the
   reference aborts otherwise, nln/mac.c:734).
4. The emitted calls, pQ1 flushes and leaf remainders, evaluated by the FP64
   oracle kernels, give the direct sum's forces to the tree's accuracy; every
   emitted record equals the one its append names.
"""
import numpy as np
import pytest

import oracle
from oracle import dlmac

pytest.importorskip("pgrav")

MXN = dict(min_hsrc=16, min_qsrc=16, min_hsink=8, min_qsink=8)


def _walk(n=2048, seed=1, gated=1, log=1, **kw):
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.plummer(n, seed=seed)
    dl = dict(gated=gated, log_calls=log, **MXN)
    return pos, mass, DLWalk(pos, mass, dl=dl, **kw)


def _offsets(W):
    nimg = W.stats()["nimage"]
    f = round(nimg ** (1 / 3))
    h = f // 2
    box = W.params.box
    off = np.zeros((nimg, 3), np.float32)
    for i in range(-h, h + 1):
        for j in range(-h, h + 1):
            for k in range(-h, h + 1):
                off[(i + h) * f * f + (j + h) * f + (k + h)] = (i * box, j * box, k * box)
    return off


def test_dl_mac_known_answers():
    # node 0: a cell of 30 bodies (hexa), node 1: a cell of 10 (quad), node 2: 3 bodies
    cells = np.array([[3.0, 10, 0, 0, 2.0, 1.0, 1.5, 1.2],
                      [1.0, 5, 0, 0, 2.0, 0.5, 1.0, 1.0],
                      [0.3, 0, 3, 0, 0.8, 0.2, 0.8, 0.8]], np.float32)
    info = np.array([[0, 30, 1, 1], [30, 10, 1, 2], [40, 3, 1, 3]], np.int64)
    bodies = np.zeros((43, 4), np.float32)
    bodies[:, 0] = 0.1
    off = np.zeros((1, 3), np.float32)
    kw = dict(cells=cells, node_info=info, bodies=bodies, offsets=off, dlfac=6.0, p2cut=7,
              p4cut=20, gated=False, leaf_max_n=56)
    sink = dict(pos=(0, 0, 0), bmax=0.1, isbody=False)
    # node 0 at distance 10: (2 + 1)^2 = 9 < 100 -> monopole
    r, a, n = dlmac.dl_mac(sink, [(0, 0)], **kw)
    assert r == [0] and a == [(0, 0, 0)] and n == 30
    # at distance 2.6: fails M (9), fails Q ((1.5+1)^2 = 6.25 < 6.76 passes!) -> Q
    c2 = cells.copy()
    c2[0, 1] = 2.6
    r, a, n = dlmac.dl_mac(sink, [(0, 0)], **dict(kw, cells=c2))
    assert r == [0] and a == [(1, 0, 0)]
    # at distance 2.4: fails M and Q (6.25 > 5.76), passes H ((1.2+1)^2 = 4.84) -> H
    c2[0, 1] = 2.4
    r, a, n = dlmac.dl_mac(sink, [(0, 0)], **dict(kw, cells=c2))
    assert r == [0] and a == [(2, 0, 0)]
    # at distance 1: nothing passes; smallest rcrit (hexa) 1.2 > 6 * 0.1 -> split source,
    # terminal expansion: its 30 bodies in order
    c2[0, 1] = 1.0
    r, a, n = dlmac.dl_mac(sink, [(0, 0)], **dict(kw, cells=c2))
    assert r == [0] and a == [(0, -(b + 1), 0) for b in range(30)] and n == 30
    # ... gated with leaf_max_n 20: stays split source
    r, a, n = dlmac.dl_mac(sink, [(0, 0)], **dict(kw, cells=c2, gated=True, leaf_max_n=20))
    assert r == [2] and a == []
    # a bigger sink: 6 * 0.25 = 1.5 > 1.2 -> split sink
    r, a, n = dlmac.dl_mac(dict(sink, bmax=0.25), [(0, 0)], **dict(kw, cells=c2))
    assert r == [1] and a == []
    # the quad cell (10 bodies, no hexa) at distance 1: fails M (6.25) and Q (2.25);
    # its smallest radius is rcrit_q = 1.0
    c2[1, 1] = 1.0
    r, a, n = dlmac.dl_mac(dict(sink, bmax=0.17), [(1, 0)], **dict(kw, cells=c2))
    assert r == [1]  # 1.02 > 1.0
    r, a, n = dlmac.dl_mac(dict(sink, bmax=0.16), [(1, 0)], **dict(kw, cells=c2))
    assert r == [0] and len(a) == 10  # 0.96 <= 1.0: split source, expanded
    # node 2 (3 bodies, monopole only) at distance 3: (0.8 + 0.2)^2 = 1 < 9 -> M
    r, a, n = dlmac.dl_mac(sink, [(2, 0)], **kw)
    assert r == [0] and a == [(0, 2, 0)]
    # a body source waits for a body sink
    r, a, n = dlmac.dl_mac(sink, [(-5, 0)], **kw)
    assert r == [1] and a == []
    r, a, n = dlmac.dl_mac(dict(sink, isbody=True, bmax=0.0), [(-5, 0), (2, 0)], **kw)
    assert r == [0, 0] and a == [(0, -5, 0), (0, 2, 0)] and n == 4
    # unequal masses in a terminal expansion are an error, as in the reference
    b2 = bodies.copy()
    b2[3, 0] = 0.2
    with pytest.raises(ValueError):
        dlmac.dl_mac(sink, [(0, 0)], **dict(kw, cells=c2, bodies=b2))


@pytest.mark.parametrize("gated,mac,order", [(1, "theta", 4), (0, "theta", 4),
                                             (1, "tolerance", 4), (1, "theta", 2),
                                             (1, "tolerance", 2)])
def test_dl_walk_calls_match_oracle(gated, mac, order):
    pos, mass, W = _walk(n=1024 if gated else 512, gated=gated, order=order, mac=mac,
                         theta=0.6, errtol=2e-6, frac_tol=0.0)
    cells, info, bodies = W.cells, W.node_info, W.bodies
    off = _offsets(W)
    p = W.params
    p4 = p.hexa_ncut if order >= 4 else 0
    p2 = p.quad_ncut if order >= 2 else 0
    calls = W.mac_calls()
    assert len(calls) == W.stats()["mac_calls"] > 0
    nacc = [0, 0, 0]
    for sink_ent, srcs, apps in calls:
        if sink_ent >= 0:
            sink = dict(pos=cells[sink_ent, 1:4], bmax=cells[sink_ent, 5], isbody=False)
        else:
            sink = dict(pos=bodies[-sink_ent - 1, 1:4], bmax=0.0, isbody=True)
        res, app, _ = dlmac.dl_mac(sink, [(s[0], s[1]) for s in srcs], cells, info, bodies, off,
                                   W.dl.dlfac, p2, p4, bool(gated), W.dl.leaf_max_n)
        assert list(srcs[:, 2]) == res, sink_ent
        assert [tuple(a) for a in apps.tolist()] == app, sink_ent
        for k, _, _ in app:
            nacc[k] += 1
    # (the ungated walk expands the root itself at the first small sink: bodies only)
    assert nacc[0] > 0 and (order < 4 or not gated or nacc[2] > 0)


@pytest.mark.parametrize("periodic", [0, 2])
def test_dl_walk_interactions(periodic):
    kw = dict(periodic="images", box=1.0, nimage=1) if periodic else {}
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.plummer(1024, seed=3)
    if periodic:
        pos = (pos - pos.min(0)) / (np.ptp(pos, 0).max() * 1.01)
    W = DLWalk(pos, mass, dl=dict(gated=1, **MXN), order=4, theta=0.6, **kw)
    st = W.stats()
    assert st["bad_interactions"] == 0
    np.testing.assert_array_equal(W.interactions, st["nimage"] * st["n"])


def _evaluate(W, eps2):
    """FP64 oracle evaluation of everything emitted: the M x N calls (pH / pQ,
    unsoftened, as the GPU kernels), the pQ1 flushes and the leaves (the
    reference's CPU kernels, softened)."""
    body = W.bodies
    n = body.shape[0]
    acc = np.zeros((n, 4))
    src = W.sources
    for routine, first, m, so, sn, sz in W.calls:
        order = 4 if routine == 0 else 2
        acc[first:first + m] += oracle.pmn(order, body[first:first + m, 1:4], src[so:so + sn * sz],
                                           sn, sz)
    qseg = W.qseg
    for base, m, s0, so, slen, _ in W.qflush:
        seg = qseg[s0:s0 + m]
        acc[base:base + m] += oracle.pq1_mnss(body[base:base + m, 1:4], seg,
                                              src[so:so + slen * 13])
    for j in range(n):
        ho, hn, qo, qn, mo, mn = W.tails[j]
        p = body[j:j + 1, 1:4]
        if hn:
            acc[j] += oracle.pmn(4, p, src[ho:ho + hn * 29], hn, 29, eps2=eps2)[0]
        if qn:
            acc[j] += oracle.pmn(2, p, src[qo:qo + qn * 13], qn, 13, eps2=eps2)[0]
        if mn:
            acc[j] += oracle.pmn(0, p, src[mo:mo + mn * 4], mn, 4, eps2=eps2)[0]
    return acc


@pytest.mark.parametrize("mac", ["theta", "tolerance"])
def test_dl_walk_forces_match_direct_sum(mac):
    eps = 1e-3
    pos, mass, W = _walk(n=4096 if mac == "theta" else 2048, log=0, order=4, mac=mac, theta=0.6,
                         errtol=1e-5, frac_tol=0.0)
    st = W.stats()
    if mac == "theta":  # (at this N the tolerance radii leave no sink M x N work)
        assert st["ph_calls"] > 0 and st["pq_calls"] + st["pq1_flushes"] > 0, st
    acc = _evaluate(W, eps * eps)
    b = W.bodies.astype(np.float64)
    ref = oracle.direct(b[:, 1:4], b[:, 1:4], b[:, 0], soft_type=1, eps=eps)
    da = np.linalg.norm(acc[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    assert np.median(da) < 2e-4 and da.max() < 2e-2, (np.median(da), da.max())
    dp = np.abs(acc[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
    assert np.median(dp) < 2e-4, np.median(dp)


def test_dl_walk_emitted_records():
    """Each body's leaf Mvec remainder holds, in order, exactly the M records its
    sink path appended (the logged appends along the path, re-created by the
    oracle from the cells and bodies), padded with zero records to 8."""
    pos, mass, W = _walk(n=512, order=4, theta=0.6)
    cells, bodies, recs = W.cells, W.bodies, W.records
    off = _offsets(W)
    calls = W.mac_calls()
    # the path of every sink: the walk is depth first, so a body's ancestors' calls
    # are the most recent calls on the stack of open sinks
    child_off, children = W.child_off, W.children
    parent = {}
    for c in range(child_off.size - 1):
        for e in children[child_off[c]:child_off[c + 1]]:
            parent[int(e)] = c
    by_sink = {}
    for sink_ent, srcs, apps in calls:
        by_sink.setdefault(sink_ent, []).append(apps)
    src = W.sources
    rng = np.random.default_rng(0)
    for j in rng.choice(bodies.shape[0], 40, replace=False):
        path, e = [], -(int(j) + 1)
        while True:
            path.append(e)
            if e not in parent:
                break
            e = parent[e]
        want = []
        for e in reversed(path):
            for apps in by_sink.get(e, []):
                want += [tuple(a) for a in apps.tolist() if a[0] == dlmac.M]
        ho, hn, qo, qn, mo, mn = W.tails[j]
        got = dlmac.soa_records(src[mo:mo + mn * 4], mn, 4)
        assert mn == (len(want) + 7) // 8 * 8
        exp = np.array([dlmac.record(k, ent, img, cells, recs, bodies, off) for k, ent, img in want])
        np.testing.assert_array_equal(got[:len(want)], exp)
        assert not got[len(want):].any()


@pytest.mark.parametrize("sb", [0, 1])
def test_tail_spans_equal_copied_tails(sb):
    """tail_spans = 1 emits every sink's new stack blocks once and gives each body's
    leaf remainders as runs of them: per body and list, the records the runs hold
    are exactly the copied remainders of tail_spans = 0, and everything else the
    walk emits is unchanged."""
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    if sb:
        pos, mass = ics.lattice_grf(10, seed=4, rms=0.3)
        kw = dict(periodic="images", box=1.0, nimage=1)
        dl = dict(subtract_background=1, geometric_center=1, **MXN)
    else:
        pos, mass = ics.plummer(2048, seed=5)
        kw, dl = {}, dict(MXN)
    A = DLWalk(pos, mass, dl=dl, order=4, theta=0.6, **kw)
    B = DLWalk(pos, mass, dl=dict(dl, tail_spans=1), order=4, theta=0.6, **kw)
    np.testing.assert_array_equal(A.calls[:, :3], B.calls[:, :3])
    np.testing.assert_array_equal(A.interactions, B.interactions)
    np.testing.assert_array_equal(A.tails[:, 1::2], B.tails[:, 1::2])
    assert (B.tails[:, 0::2] == -1).all()
    sa, sbf = A.sources, B.sources
    for k in range(A.calls.shape[0]):
        _, _, _, oa, n, sz = A.calls[k]
        ob = B.calls[k, 3]
        np.testing.assert_array_equal(sa[oa:oa + n * sz], sbf[ob:ob + n * sz])
    sizes = (29, 13, 4)
    for j in range(A.bodies.shape[0]):
        ra, rb = A.leaf_runs(j), B.leaf_runs(j)
        for kind in range(3):
            sz = sizes[kind]
            want = np.concatenate([sa[o:o + n * sz] for o, n in ra[kind]] or [np.zeros(0, np.float32)])
            got = np.concatenate([sbf[o:o + n * sz] for o, n in rb[kind]] or [np.zeros(0, np.float32)])
            np.testing.assert_array_equal(got, want, err_msg="body %d list %d" % (j, kind))
    assert B.stats()["source_floats"] < A.stats()["source_floats"]
