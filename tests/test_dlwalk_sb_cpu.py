"""The subtract_background DL walk (DLRcritMACsb, nln/mac.c:1565-1771) on the CPU.

1. The oracle's pieces against what the mathematics fixes: mpole_add_cube's
   constants (nln/cofm.c:626-648) against Gauss-Legendre integration of the cube's
   moments (exact for polynomials); the cube's field against an independent
   semi-analytic quadrature, its far field (a point mass), its odd symmetry and
   Poisson's equation (div g = -4 pi inside, 0 outside).
2. The library's subtracted records against records built by the oracle from the
   cell's bodies plus a Gauss-Legendre point set of the negative cube; the
   unsubtracted umass / uxxxx / uyyyy / uxxyy against the plain walk's records.
3. Every MAC call the library makes, replayed through the oracle: same results,
   background flags, appends (anti-masses, MMvec segments) and near counts.
4. Every body accounts for nimage N bodies; every MMvec segment appended on a
   body's path is evaluated exactly once for it (pMM_mnss_sK1 queue, host
   mxn_mono runs or its leaf), and the queue's packing is consistent.
5. The leaf's near-cube term equals -rho0 times the oracle's cube field.
"""
import itertools

import numpy as np
import pytest

import oracle
from oracle import sbmac

pytest.importorskip("pgrav")

GL3 = np.polynomial.legendre.leggauss(3)


def _walk(ng=16, seed=4, nimage=1, log=1, order=4, mac="theta", clustered=0, **dl):
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    if clustered:
        # halos and a background in the unit box: sparse bodies next to dense sinks
        pos, _ = ics.clustered(clustered, seed=seed)
        pos = (pos - pos.min(0)) / (np.ptp(pos, 0).max() * 1.001)
        mass = np.full(clustered, 1.0 / clustered)
    else:
        pos, mass = ics.lattice_grf(ng, seed=seed)
    d = dict(subtract_background=1, log_calls=log, min_hsrc=16, min_qsrc=16, min_hsink=8,
             min_qsink=8, min_msrc=16)
    d.update(dl)
    W = DLWalk(pos, mass, dl=d, order=order, mac=mac, theta=0.6, errtol=2e-6, frac_tol=0.0,
               periodic="images", box=1.0, nimage=nimage)
    return pos, mass, W


def _offsets(W):
    nimg = W.stats()["nimage"]
    f = round(nimg ** (1 / 3))
    h = f // 2
    off = np.zeros((nimg, 3), np.float32)
    for i, j, k in itertools.product(range(-h, h + 1), repeat=3):
        off[(i + h) * f * f + (j + h) * f + (k + h)] = (i, j, k)
    return off * np.float32(W.params.box)


def _rho0(W, mass):
    # the library sums the caller's FP64 masses in its body order
    m0 = np.asarray(mass, np.float64)[W.perm].sum()
    r0 = float(np.float32(0.5 * W.params.box))
    return m0 / (2.0 * r0) ** 3, r0


# ---------------------------------------------------------------- the oracle's pieces

def test_mpole_add_cube_constants():
    """The cube's raw moments about its centre by the 3-point Gauss-Legendre product
    rule (exact to degree 5 per axis) equal mpole_add_cube's, for every monomial of
    degree <= 4 (the odd and mixed ones vanish)."""
    a, rho = 0.37, -2.5
    x, w = GL3
    pts = 0.5 * a * x
    wt = 0.5 * a * w
    got = sbmac.mpole_add_cube(a, rho, {})
    for i, j, k in itertools.product(range(5), repeat=3):
        if i + j + k > 4:
            continue
        m = rho * sum(wt[p] * wt[q] * wt[s] * pts[p] ** i * pts[q] ** j * pts[s] ** k
                      for p in range(3) for q in range(3) for s in range(3))
        assert got.get((i, j, k), 0.0) == pytest.approx(m, rel=1e-13, abs=1e-15), (i, j, k)


def test_cube_field_closed_form_pinned():
    rng = np.random.default_rng(0)
    a = 0.3
    pts = [rng.uniform(-3, 3, 3) for _ in range(12)]            # mostly outside
    pts += [rng.uniform(-0.29, 0.29, 3) for _ in range(8)]      # inside
    pts += [np.array([0.3, 0.1, -0.2]), np.array([0.0, 0.0, 0.7]), np.array([0.3, 0.3, 0.3])]
    for r in pts:
        g = sbmac.cube_field(r, a)
        q = sbmac.cube_field_quad(r, a)
        np.testing.assert_allclose(g, q, rtol=1e-9, atol=1e-11 * np.abs(q).max())
    # far away: a point mass (2a)^3 at the centre, to (a / r)^4
    for d in (np.array([40.0, 3.0, -7.0]), np.array([-25.0, 31.0, 12.0])):
        M = (2 * a) ** 3
        pm = -M * d / np.linalg.norm(d) ** 3
        np.testing.assert_allclose(sbmac.cube_field(d, a), pm, rtol=1e-5)
    # the centre, odd symmetry, and the linear field near the centre, -(4 pi / 3) r
    np.testing.assert_allclose(sbmac.cube_field([0, 0, 0], a), 0, atol=1e-14)
    for r in pts:
        np.testing.assert_allclose(sbmac.cube_field(-r, a), -sbmac.cube_field(r, a),
                                   atol=1e-12)
    h = np.array([1e-4, -2e-4, 5e-5])
    np.testing.assert_allclose(sbmac.cube_field(h, a), -4 * np.pi / 3 * h, rtol=1e-6)
    # Poisson: div g = -4 pi rho
    eps = 1e-5
    for r, rho in ((np.array([0.1, -0.05, 0.2]), 1.0), (np.array([0.5, 0.2, -0.4]), 0.0)):
        div = sum((sbmac.cube_field(r + eps * e, a)[k] - sbmac.cube_field(r - eps * e, a)[k])
                  / (2 * eps) for k, e in enumerate(np.eye(3)))
        assert div == pytest.approx(-4 * np.pi * rho, abs=1e-5)


# ---------------------------------------------------------------- the library

def test_sb_records_and_unsubtracted_fields():
    pos, mass, W = _walk(ng=12, log=0)
    _, _, P = _walk(ng=12, log=0, subtract_background=0)
    rec, urec, ubs = W.records, P.records, W.ubs
    info, geom, bodies, cells = W.node_info, W.node_geom, W.bodies, W.cells
    rho0, _ = _rho0(W, mass)
    x, w = GL3
    np.testing.assert_array_equal(W.cells, P.cells)
    np.testing.assert_array_equal(ubs[:, 0], urec[:, 0])
    np.testing.assert_array_equal(ubs[:, 1:], urec[:, [20, 24, 22]])
    checked = 0
    for c in np.flatnonzero(info[:, 1] >= 20)[:40]:
        b0, nb = int(info[c, 0]), int(info[c, 1])
        e = float(geom[c, 3])
        cen = cells[c, 1:4].astype(np.float64)
        # the cube as 27 point masses about the centre of mass (mpole_add_cube takes
        # the cube's moments about the expansion centre)
        g = np.array(list(itertools.product(range(3), repeat=3)))
        cpts = cen + 0.5 * e * x[g]
        cw = -rho0 * (0.5 * e) ** 3 * np.prod(w[g], axis=1)
        P64 = np.vstack([W.bodies[b0:b0 + nb, 1:4].astype(np.float64), cpts])
        m64 = np.concatenate([np.full(nb, bodies[b0, 0], np.float64), cw])
        # the library's centre is the FP64 centre of mass; the float one is close
        want = oracle.cell_record(P64, m64, cen, float(cells[c, 5]))
        got = rec[c].astype(np.float64)
        want[5:8] = 0.0
        scale = np.abs(want).max()
        # (the subtracted mass nearly cancels: compare it on the unsubtracted scale)
        assert abs(got[0] - want[0]) <= 2e-6 * ubs[c, 0]
        assert got[4] == pytest.approx(want[4], rel=1e-6)
        np.testing.assert_allclose(got[8:], want[8:], atol=2e-4 * scale)
        # the moments the cube leaves alone are the plain record's
        np.testing.assert_array_equal(rec[c, 8:20], urec[c, 8:20])
        np.testing.assert_array_equal(rec[c, [21, 23, 25, 26, 27, 28]],
                                      urec[c, [21, 23, 25, 26, 27, 28]])
        checked += 1
    assert checked > 10


@pytest.mark.parametrize("ng,order,mac,nimage,dl", [
    (16, 4, "theta", 1, {}),
    (12, 4, "tolerance", 1, {}),
    (12, 2, "theta", 1, {}),
    (10, 4, "theta", 2, {}),
    (12, 4, "theta", 1, dict(do_pM=0)),
    (16, 4, "theta", 1, dict(leaf_max_n=8, dlfac=2.0)),
    # clustered: anti-masses of bodies and split cells, host mxn_mono runs; with the
    # near cubes of a larger run (nx 512) also the empty-octant anti-masses
    (4096, 4, "theta", 1, dict(clustered=1)),
    (8192, 4, "theta", 1, dict(clustered=1, nx=512.0)),
    (4096, 4, "tolerance", 1, dict(clustered=1, nx=256.0)),
    # cells expanded about their key cube's centre
    (12, 4, "tolerance", 1, dict(geometric_center=1)),
    (4096, 4, "theta", 1, dict(clustered=1, nx=512.0, geometric_center=1)),
])
def test_sb_walk_calls_match_oracle(ng, order, mac, nimage, dl):
    dl = dict(dl)
    if dl.pop("clustered", 0):
        dl["clustered"] = ng
    pos, mass, W = _walk(ng=ng, order=order, mac=mac, nimage=nimage, **dl)
    cells, info, bodies, geom, bl = W.cells, W.node_info, W.bodies, W.node_geom, W.body_leaf
    off = _offsets(W)
    rho0, r0 = _rho0(W, mass)
    uc = sbmac.ucell(r0, rho0)
    p = W.params
    p4 = p.hexa_ncut if order >= 4 else 0
    p2 = p.quad_ncut if order >= 2 else 0
    calls = W.mac_calls()
    flags, lsb = W.log_flags, W.log_sb
    assert len(calls) == lsb.shape[0] == W.stats()["mac_calls"] > 0
    src_at = 0
    kinds = {}
    for k, (sink_ent, srcs, apps) in enumerate(calls):
        fl = flags[src_at:src_at + len(srcs)]
        src_at += len(srcs)
        if sink_ent >= 0:
            sp, sbm, nd = cells[sink_ent, 1:4], cells[sink_ent, 5], int(info[sink_ent, 1])
        else:
            sp, sbm, nd = bodies[-sink_ent - 1, 1:4], 0.0, 1
        st = lsb[k]
        sink = dict(pos=sp, bmax=sbm, daughters=nd, cen=st[:3], cr=st[3], cr2=st[4],
                    clevel=int(st[5]), done=bool(st[6]))
        res, fout, app, near, _ = sbmac.dl_mac_sb(
            sink, [(s[0], s[1], f) for s, f in zip(srcs, fl[:, 0])], cells, info, geom, bl,
            bodies, off, uc, rho0, p.box, W.dl.dlfac, p2, p4, W.dl.leaf_max_n, W.dl.do_pM)
        assert list(srcs[:, 2]) == res, (k, sink_ent)
        assert list(fl[:, 1]) == fout, (k, sink_ent)
        assert [tuple(a) for a in apps.tolist()] == app, (k, sink_ent)
        for a in app:
            kinds[a[0]] = kinds.get(a[0], 0) + 1
    assert kinds.get(sbmac.MM, 0) > 0 if W.dl.do_pM else kinds.get(sbmac.MM, 0) == 0
    st = W.sb_stats()
    if "clustered" in dl:
        assert st["anti_cell"] > 0 and st["host_mono"] > 0, st
        if "nx" in dl:
            assert st["anti_body"] > 0 and st["anti_empty"] > 0, st
    assert st["mm_segments"] == kinds.get(sbmac.MM, 0)
    assert st["anti_body"] == kinds.get(sbmac.ANTI_BODY, 0)
    assert st["anti_cell"] == kinds.get(sbmac.ANTI_CELL, 0)
    assert st["anti_empty"] == sum(v for c, v in kinds.items()
                                   if sbmac.ANTI_EMPTY <= c < sbmac.ANTI_EMPTY + 8)


def test_sb_interactions_and_mm_accounting():
    pos, mass, W = _walk(ng=16)
    st = W.stats()
    assert st["bad_interactions"] == 0
    np.testing.assert_array_equal(W.interactions, st["nimage"] * st["n"])
    n = st["n"]
    mseg = W.mseg
    # pairs per body from each route
    pairs = np.zeros(n, np.int64)
    idx_all, ssg, srcn = W.mss_index, W.mss_seg, W.mss_source_n
    for base, m, i0, e0, ss_len, g0, seg_n, _ in W.mflush:
        idx = idx_all[i0:i0 + m]
        assert idx.max() < ss_len
        for s in range(m):
            b, ln = ssg[e0 + idx[s]]
            segs = mseg[g0 + b:g0 + b + ln]
            assert b + ln <= seg_n
            cnt = int(segs[:, 1].sum()) if ln else 0
            assert cnt == srcn[e0 + idx[s]]
            pairs[base + s] += cnt
    for first, m, s0, ns, nsrc, _ in W.mcalls:
        segs = mseg[s0:s0 + ns]
        assert segs[:, 1].sum() == nsrc
        pairs[first:first + m] += nsrc
    for j in range(n):
        s0, ns, nsrc = W.mmtails[j]
        assert mseg[s0:s0 + ns, 1].sum() == nsrc
        pairs[j] += nsrc
    # ... equals the MMvec bodies appended along each body's sink path
    child_off, children = W.child_off, W.children
    parent = {}
    for c in range(child_off.size - 1):
        for e in children[child_off[c]:child_off[c + 1]]:
            parent[int(e)] = c
    per_sink = {}
    info = W.node_info
    for sink_ent, srcs, apps in W.mac_calls():
        v = sum(int(info[e, 1]) for code, e, _ in apps.tolist() if code == sbmac.MM)
        per_sink[sink_ent] = per_sink.get(sink_ent, 0) + v
    rng = np.random.default_rng(1)
    for j in rng.choice(n, 64, replace=False):
        e, want = -(int(j) + 1), 0
        while True:
            want += per_sink.get(e, 0)
            if e not in parent:
                break
            e = parent[e]
        assert pairs[j] == want, j
    assert W.sb_stats()["mns_flushes"] >= 1 and len(W.mflush) >= 1


def test_sb_background_term():
    pos, mass, W = _walk(ng=12, log=0)
    rho0, _ = _rho0(W, mass)
    body, q, bg = W.bodies, W.sb_sink, W.background
    rng = np.random.default_rng(2)
    for j in rng.choice(body.shape[0], 40, replace=False):
        r = (body[j, 1:4] - q[j, :3]).astype(np.float64)
        want = -rho0 * sbmac.cube_field(r, float(q[j, 3]))
        np.testing.assert_allclose(bg[j], want, rtol=1e-12, atol=1e-14)
    # a body's near cube is 3 x 3 x 3 of its level's cells: 27 near cells when the
    # lattice is regular enough
    assert np.median(q[:, 5]) >= 27


def test_sb_node_geometry():
    """Each node's key cube holds its bodies, its edge halves per level, and its
    octant mask has bit j (x = j & 1, y = j >> 1 & 1, z = j >> 2, CellCorner's key
    digit) set exactly where a body lies -- from the bodies' own positions."""
    pos, mass, W = _walk(clustered=4096, log=0)
    info, geom, body = W.node_info, W.node_geom, W.bodies.astype(np.float64)
    for c in range(info.shape[0]):
        b0, nb, kids, lev = (int(v) for v in info[c])
        corner, edge = geom[c, :3].astype(np.float64), float(geom[c, 3])
        assert edge == W.params.box / 2 ** lev
        p = body[b0:b0 + nb, 1:4]
        assert np.all(p >= corner - 1e-7) and np.all(p < corner + edge + 1e-7), c
        if kids:
            o = (p >= corner + 0.5 * edge).astype(int)
            mask = np.bitwise_or.reduce(1 << (o[:, 0] | o[:, 1] << 1 | o[:, 2] << 2))
            assert int(geom[c, 4]) == mask, c
        else:
            assert int(geom[c, 4]) == 0xFF
    # each body's leaf node holds it
    bl = W.body_leaf
    for j in range(0, body.shape[0], 37):
        b0, nb = int(info[bl[j], 0]), int(info[bl[j], 1])
        assert b0 <= j < b0 + nb


# ---------------------------------------------------------------- the whole scheme

def _geo_walk(ng, rms, nimage, log=1):
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.lattice_grf(ng, seed=4, rms=rms)
    W = DLWalk(pos, mass, dl=dict(subtract_background=1, geometric_center=1, log_calls=log),
               order=4, theta=0.6, periodic="images", box=1.0, nimage=nimage)
    return pos, mass, W


def _pts_field(p, q, m):
    d = q - p
    r2 = (d * d).sum(1)
    ok = r2 > 0
    return (m[ok, None] * d[ok] / (r2[ok] ** 1.5)[:, None]).sum(0)


def test_sb_lists_partition_bodies_less_background():
    """Every list entry of a body's path evaluated exactly -- a cell's bodies (less
    its key cube of density rho0 where the record carries the subtraction), the
    anti-masses as the cubes they stand for, MMvec bodies -- plus the leaf's near
    cube and the subtract_background far field is the Ewald sum (nln/ewald.c,
    oracle): the walk covers bodies less background over the explicit images once
    each.  With geometric_center the signed mass of each body's lists plus its near
    cube is zero and every near cube holds 27 cells (the check the reference keeps
    disabled, nln/mac.c:697-707)."""
    from pgrav import batch
    pos, mass, W = _geo_walk(12, 0.02, 2)
    cells, info, geom, body = W.cells, W.node_info, W.node_geom, W.bodies.astype(np.float64)
    rho0 = mass.sum()
    off = _offsets(W).astype(np.float64)
    assert np.all(W.sb_sink[:, 5] == 27)
    parent = {}
    co, ch = W.child_off, W.children
    for c in range(co.size - 1):
        for e in ch[co[c]:co[c + 1]]:
            parent[int(e)] = c
    apps = {}
    for sink, _, ap in W.mac_calls():
        apps.setdefault(sink, []).extend(ap.tolist())

    def cube(p, corner, edge, o):
        return -rho0 * sbmac.cube_field(p - (corner + 0.5 * edge + o), 0.5 * edge)
    rng = np.random.default_rng(0)
    sample = rng.choice(body.shape[0], 16, replace=False)
    perm = W.perm
    far = batch.ewald_le_host(pos, mass, 1.0, 2, subtract_background=True)
    ex = oracle.direct_ewald(pos[perm[sample]], pos, mass, 1.0)
    errs = []
    for si, j in enumerate(sample):
        p = body[j, 1:4]
        e, acc, msum = -(int(j) + 1), np.zeros(3), 0.0
        path = []
        while True:
            path.append(e)
            if e not in parent:
                break
            e = parent[e]
        for s in path:
            for code, ent, img in apps.get(s, []):
                o = off[img]
                if ent < 0:
                    b = -ent - 1
                    assert code == sbmac.M  # no body anti-mass in a lattice this regular
                    acc += _pts_field(p, body[b:b + 1, 1:4] + o, body[b:b + 1, 0])
                    msum += body[b, 0]
                    continue
                b0, nb = int(info[ent, 0]), int(info[ent, 1])
                corner, edge = geom[ent, :3].astype(np.float64), float(geom[ent, 3])
                if code in (sbmac.Q, sbmac.H, sbmac.QB, sbmac.HB, sbmac.MM):
                    acc += _pts_field(p, body[b0:b0 + nb, 1:4] + o, body[b0:b0 + nb, 0])
                    msum += body[b0:b0 + nb, 0].sum()
                    if code in (sbmac.Q, sbmac.H):
                        acc += cube(p, corner, edge, o)
                        msum -= rho0 * edge ** 3
                elif code == sbmac.ANTI_CELL:
                    acc += cube(p, corner, edge, o)
                    msum -= rho0 * edge ** 3
                else:
                    raise AssertionError("unexpected append %d" % code)
        cr = float(W.sb_sink[j, 3])
        assert abs(msum - rho0 * (2 * cr) ** 3) < 1e-6 * rho0
        acc += W.background[j] + far[perm[j], :3]
        errs.append(np.linalg.norm(acc - ex[si, :3]))
    rms = np.sqrt((ex[:, :3] ** 2).sum(1).mean())
    errs = np.array(errs)
    assert np.median(errs) < 1e-4 * rms and errs.max() < 1e-3 * rms, (np.median(errs) / rms,
                                                                       errs.max() / rms)


def _evaluate_sb(W, eps_inv):
    """FP64 oracle evaluation of everything the sb walk emits: the M x N calls, the
    pQ1 and pMM_mnss_sK1 flushes, the host mxn_mono runs, the leaves (K1-softened
    M lists, as smooth_type 4 with MMvec, nln/mac.c:221-224) and the near cube."""
    body = W.bodies
    n = body.shape[0]
    acc = np.zeros((n, 4))
    src = W.sources
    bp = body[:, 1:4]
    m0 = float(body[0, 0])
    for routine, first, m, so, sn, sz in W.calls:
        acc[first:first + m] += oracle.pmn(4 if routine == 0 else 2, bp[first:first + m],
                                           src[so:so + sn * sz], sn, sz)
    for base, m, s0, so, slen, _ in W.qflush:
        acc[base:base + m] += oracle.pq1_mnss(bp[base:base + m], W.qseg[s0:s0 + m],
                                              src[so:so + slen * 13])
    for base, m, i0, e0, ss_len, g0, seg_n, _ in W.mflush:
        out, bad = oracle.pmm1_mnss_sk1(bp[base:base + m], W.mss_index[i0:i0 + m],
                                        W.mss_seg[e0:e0 + ss_len], W.mseg[g0:g0 + seg_n], bp,
                                        W.mss_source_n[e0:e0 + ss_len], m0, eps_inv)
        assert bad == 0
        acc[base:base + m] += out
    for first, m, s0, ns, nsrc, _ in W.mcalls:
        acc[first:first + m] += oracle.pmm_sk1(bp[first:first + m], W.mseg[s0:s0 + ns], bp, m0,
                                               eps_inv)
    for j in range(n):
        ho, hn, qo, qn, mo, mn = W.tails[j]
        p = bp[j:j + 1]
        if hn:
            acc[j] += oracle.pmn(4, p, src[ho:ho + hn * 29], hn, 29)[0]
        if qn:
            acc[j] += oracle.pmn(2, p, src[qo:qo + qn * 13], qn, 13)[0]
        if mn:
            acc[j] += oracle.pm_sk1(p, src[mo:mo + mn * 4], mn, eps_inv)[0]
        s0, ns, nsrc = W.mmtails[j]
        if ns:
            acc[j] += oracle.pmm_sk1(p, W.mseg[s0:s0 + ns], bp, m0, eps_inv)[0]
    acc[:, :3] += W.background
    return acc


@pytest.mark.parametrize("rms,bound", [(0.02, 3e-2), (0.3, 1e-2)])
def test_sb_forces_match_ewald(rms, bound):
    """The emitted lists through the oracle kernels plus the far field against the
    Ewald sum.  What remains is the scheme's own approximation: the subtracted
    cells' truncated expansions and the split cells' background as point masses
    (the exact-entry evaluation above closes to 1e-5)."""
    from pgrav import batch
    pos, mass, W = _geo_walk(12, rms, 1, log=0)
    acc = _evaluate_sb(W, 1e5)
    perm = W.perm
    acc[:, :3] += batch.ewald_le_host(pos, mass, 1.0, 1, subtract_background=True)[perm, :3]
    rng = np.random.default_rng(0)
    s = rng.choice(pos.shape[0], 48, replace=False)
    ex = oracle.direct_ewald(pos[perm[s]], pos, mass, 1.0)
    rmsf = np.sqrt((ex[:, :3] ** 2).sum(1).mean())
    e = np.linalg.norm(acc[s, :3] - ex[:, :3], axis=1) / rmsf
    assert np.median(e) < bound, np.median(e)


def test_geometric_center_records():
    """geometric_center: each cell's record is its bodies' expansion about the key
    cube's centre (dipole included) and bmax bounds its bodies' distance from that
    centre within the cube's half-diagonal (nln/cofm.c:141-146, :183-222)."""
    pos, mass, W = _walk(clustered=4096, log=0, subtract_background=0, geometric_center=1)
    info, geom, cells, rec, body = W.node_info, W.node_geom, W.cells, W.records, W.bodies
    checked = 0
    for c in range(info.shape[0]):
        b0, nb = int(info[c, 0]), int(info[c, 1])
        if nb < 2:
            continue
        edge = float(geom[c, 3])
        cen = geom[c, :3].astype(np.float64) + 0.5 * edge
        np.testing.assert_allclose(cells[c, 1:4], cen, rtol=0, atol=1e-7)
        p = body[b0:b0 + nb, 1:4].astype(np.float64)
        dmax = np.sqrt(((p - cen) ** 2).sum(1)).max()
        # (the float bodies sit within 1e-7 of the FP64 positions the library used)
        assert dmax - 1e-7 <= cells[c, 5] <= 0.5 * edge * np.sqrt(3) * (1 + 1e-6)
        if nb >= 20 and checked < 60:
            want = oracle.cell_record(p, np.full(nb, float(body[b0, 0])), cen, float(cells[c, 5]))
            got = rec[c].astype(np.float64)
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * np.abs(want).max())
            checked += 1
    assert checked > 20
