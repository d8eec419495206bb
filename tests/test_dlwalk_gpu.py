"""The drop-in ABI (nln/gravcuda.h) consuming the library's own DL-walk lists.

pgrav_dl_walk emits the reference's M x N calls -- grav_mn_CUDA("pH" / "pQ") on
SoA-8 Hvec / Qvec blocks and grav_qnss_CUDA("pQ1_qnss") flushes of the pQ1
queue -- and each body's leaf remainders.  Here every emitted call goes
through the C ABI into the body table on the GPU, exactly as the reference's
walk would issue it (WalkInitSinkCUDA -> qallocCUDA -> grav_*_CUDA ->
WalkTerminateSinkCUDA), and:

1. the GPU's additions equal the FP64 oracle's evaluation of the same emitted
   arrays to FP32 accuracy (2e-5 of each sink's largest component);
2. with the leaf remainders added (the reference evaluates those on the CPU; here
   the FP64 oracle), every body's force matches the FP64 direct sum to the
   tree's accuracy.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

STRIDE = 14  # sizeof(body)/sizeof(float), nln/physics.h:33-44
MXN = dict(min_hsrc=16, min_qsrc=16, min_hsink=8, min_qsink=8)


@pytest.mark.parametrize("mac", ["theta", "tolerance"])
def test_dropin_consumes_dl_lists(gpu_lib, mac):
    pgrav = gpu_lib
    from pgrav import ics
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.plummer(4096, seed=1)
    W = DLWalk(pos, mass, dl=MXN, order=4, mac=mac, theta=0.6, errtol=2.4e-7, frac_tol=0.0)
    st = W.stats()
    print("dl walk %s: %d pH, %d pQ calls, %d pQ1 flushes" % (mac, st["ph_calls"], st["pq_calls"],
                                                              st["pq1_flushes"]), flush=True)
    assert st["ph_calls"] > 0
    body = W.bodies
    n = body.shape[0]
    btab = np.zeros((n, STRIDE), np.float32)
    btab[:, 0] = body[:, 0]
    btab[:, 1:4] = body[:, 1:4]
    flat = btab.reshape(-1)
    src = W.sources
    g = pgrav.GravCuda()
    g.walk_init_sink(flat, STRIDE)
    ref = np.zeros((n, 4))
    for routine, first, m, so, sn, sz in W.calls:
        f = np.ascontiguousarray(src[so:so + sn * sz])
        q, _ = g.qalloc()
        g.grav_mn(("pH", "pQ")[routine], int(first), int(m), f, int(sn), int(sz), 0.0, q)
        ref[first:first + m] += oracle.pmn(4 if routine == 0 else 2, body[first:first + m, 1:4],
                                           f, sn, sz)
    qseg = W.qseg
    for base, m, s0, so, slen, _ in W.qflush:
        seg = np.ascontiguousarray(qseg[s0:s0 + m]).reshape(-1)
        buf = np.ascontiguousarray(src[so:so + slen * 13])
        q, _ = g.qalloc()
        g.grav_qnss("pQ1_qnss", int(base), int(m), seg, buf, int(slen), q)
        ref[base:base + m] += oracle.pq1_mnss(body[base:base + m, 1:4], seg, buf)
    g.walk_terminate_sink(flat, STRIDE)
    got = btab[:, 7:11].astype(np.float64)
    hit = np.abs(ref).max(axis=1) > 0
    assert hit.sum() > 0
    scale = np.abs(ref[hit]).max(axis=1, keepdims=True)
    err = np.abs(got[hit] - ref[hit]) / scale
    assert err.max() < 2e-5, err.max()
    assert np.all(got[~hit] == 0)

    # the whole force: the GPU's M x N work plus the leaves (FP64 oracle, Plummer
    # softened as the reference's CPU kernels) against the FP64 direct sum
    eps = 1e-3
    acc = got.copy()
    for j in range(n):
        ho, hn, qo, qn, mo, mn = W.tails[j]
        p = body[j:j + 1, 1:4]
        if hn:
            acc[j] += oracle.pmn(4, p, src[ho:ho + hn * 29], hn, 29, eps2=eps * eps)[0]
        if qn:
            acc[j] += oracle.pmn(2, p, src[qo:qo + qn * 13], qn, 13, eps2=eps * eps)[0]
        if mn:
            acc[j] += oracle.pmn(0, p, src[mo:mo + mn * 4], mn, 4, eps2=eps * eps)[0]
    b = body.astype(np.float64)
    dref = oracle.direct(b[:, 1:4], b[:, 1:4], b[:, 0], soft_type=1, eps=eps)
    da = np.linalg.norm(acc[:, :3] - dref[:, :3], axis=1) / np.linalg.norm(dref[:, :3], axis=1)
    assert np.median(da) < 2e-4 and da.max() < 2e-2, (np.median(da), da.max())


def test_dropin_consumes_sb_lists(gpu_lib):
    """subtract_background (DLRcritMACsb, nln/mac.c:1565-1771) with geometric_center:
    the walk's pH / pQ calls (records with dipoles and background-subtracted
    moments), its pQ1 flushes and its pMM_mnss_sK1 flushes (the grav_mns_queue
    packing, nln/mac.c:352-457) through the C ABI equal the FP64 oracle on the same
    emitted arrays; with the leaves, the near cube and the sb far field the forces
    match the Ewald sum to the scheme's accuracy."""
    pgrav = gpu_lib
    from pgrav import batch, ics
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.lattice_grf(16, seed=4, rms=0.3)
    W = DLWalk(pos, mass, dl=dict(subtract_background=1, geometric_center=1, **MXN, min_msrc=16),
               order=4, theta=0.6, periodic="images", box=1.0, nimage=1)
    st, sb = W.stats(), W.sb_stats()
    print("sb walk: %d pH, %d pQ, %d pQ1 flushes, %d pMM_mnss_sK1 flushes (%d pairs)" % (
        st["ph_calls"], st["pq_calls"], st["pq1_flushes"], sb["mns_flushes"], sb["mns_pairs"]),
        flush=True)
    assert sb["mns_flushes"] > 0 and st["ph_calls"] + st["pq1_flushes"] > 0
    body = W.bodies
    n = body.shape[0]
    btab = np.zeros((n, STRIDE), np.float32)
    btab[:, 0] = body[:, 0]
    btab[:, 1:4] = body[:, 1:4]
    flat = btab.reshape(-1)
    src = W.sources
    bp = body[:, 1:4]
    m0, eps_inv = float(body[0, 0]), 1e5
    g = pgrav.GravCuda()
    g.walk_init_sink(flat, STRIDE)
    ref = np.zeros((n, 4))
    for routine, first, m, so, sn, sz in W.calls:
        f = np.ascontiguousarray(src[so:so + sn * sz])
        q, _ = g.qalloc()
        g.grav_mn(("pH", "pQ")[routine], int(first), int(m), f, int(sn), int(sz), 0.0, q)
        ref[first:first + m] += oracle.pmn(4 if routine == 0 else 2, bp[first:first + m], f, sn, sz)
    for base, m, s0, so, slen, _ in W.qflush:
        seg = np.ascontiguousarray(W.qseg[s0:s0 + m]).reshape(-1)
        buf = np.ascontiguousarray(src[so:so + slen * 13])
        q, _ = g.qalloc()
        g.grav_qnss("pQ1_qnss", int(base), int(m), seg, buf, int(slen), q)
        ref[base:base + m] += oracle.pq1_mnss(bp[base:base + m], seg, buf)
    for base, m, i0, e0, ss_len, g0, seg_n, _ in W.mflush:
        idx = W.mss_index[i0:i0 + m]
        ssg = W.mss_seg[e0:e0 + ss_len]
        seg = W.mseg[g0:g0 + seg_n]
        srcn = W.mss_source_n[e0:e0 + ss_len]
        q, _ = g.qalloc()
        g.grav_mnss("pMM_mnss_sK1", int(base), int(m), idx, ssg.reshape(-1), seg.reshape(-1), 0,
                    srcn, m0, eps_inv, q)
        out, bad = oracle.pmm1_mnss_sk1(bp[base:base + m], idx, ssg, seg, bp, srcn, m0, eps_inv)
        assert bad == 0
        ref[base:base + m] += out
    g.walk_terminate_sink(flat, STRIDE)
    got = btab[:, 7:11].astype(np.float64)
    hit = np.abs(ref).max(axis=1) > 0
    assert hit.sum() > 0
    scale = np.abs(ref[hit]).max(axis=1, keepdims=True)
    err = np.abs(got[hit] - ref[hit]) / scale
    assert err.max() < 2e-5, err.max()
    assert np.all(got[~hit] == 0)

    # the whole force: the GPU's work, the leaves and host mxn_mono runs (the
    # reference's CPU work; the FP64 oracle here), the near cube and the far field
    acc = got.copy()
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
    perm = W.perm
    acc[:, :3] += batch.ewald_le_host(pos, mass, 1.0, 1, subtract_background=True)[perm, :3]
    rng = np.random.default_rng(0)
    s = rng.choice(n, 64, replace=False)
    ex = oracle.direct_ewald(pos[perm[s]], pos, mass, 1.0)
    rms = np.sqrt((ex[:, :3] ** 2).sum(1).mean())
    e = np.linalg.norm(acc[s, :3] - ex[:, :3], axis=1) / rms
    assert np.median(e) < 1e-2, np.median(e)
