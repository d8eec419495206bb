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
