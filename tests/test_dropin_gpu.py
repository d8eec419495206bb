"""Parity of the drop-in ABI (nln/gravcuda.h) on the GPU against the FP64 oracle.

Every test goes through the C ABI of libpgrav.so (WalkInitSinkCUDA -> qallocCUDA
-> grav_*_CUDA -> WalkTerminateSinkCUDA) and compares the accelerations and
potentials added into the body table with the oracle's FP64 restatement of the
same reference kernel.  Tolerance: FP32 evaluation, so |gpu - oracle| is
bounded by 2e-5 of the largest |component| of the sink's result (a sum of
<= a few hundred terms, each rounded at ~6e-8).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

STRIDE = 14  # sizeof(body)/sizeof(float), nln/physics.h:33-44


@pytest.fixture(params=[1, 0], ids=["deferred", "per_call"], autouse=True)
def deferral(gpu_lib, request):
    """Every test runs with the calls batched (the default) and launched one by one."""
    old = gpu_lib.set_dropin_deferral(request.param)
    yield request.param
    gpu_lib.set_dropin_deferral(old)


def _bodies(rng, n, scale=3.0):
    b = np.zeros((n, STRIDE), np.float32)
    b[:, 0] = 1.0 / n
    b[:, 1:4] = rng.uniform(-scale, scale, (n, 3))
    return b


def _cells(rng, n, spread=0.1, dipole=True):
    recs = []
    for _ in range(n):
        pos = rng.uniform(-1, 1, 3) + rng.normal(size=(30, 3)) * spread
        mass = rng.uniform(0.5, 1.5, 30) / 30
        cm = (pos * mass[:, None]).sum(0) / mass.sum()
        bmax = np.sqrt(((pos - cm) ** 2).sum(1)).max()
        r = oracle.cell_record(pos, mass, cm, bmax)
        if dipole:
            r[5:8] = rng.normal(size=3) * 0.1 * r[0]
        recs.append(r)
    return np.array(recs)


def _close(got, ref, tol=2e-5):
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(got - ref) / scale
    assert err.max() < tol, (err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("routine,order,sz", [("pH", 4, 29), ("pQ", 2, 13), ("pM", 0, 4)])
@pytest.mark.parametrize("qsign", [1, -1])
def test_grav_mn(gpu_lib, routine, order, sz, qsign):
    pgrav = gpu_lib
    rng = np.random.default_rng(100 + order)
    g = pgrav.GravCuda()
    pgrav.set_quad_phi_sign(qsign)
    try:
        btab = _bodies(rng, 300)
        btab[:, 7:11] = rng.normal(size=(300, 4))  # results are added (+=)
        before = btab[:, 7:11].copy()
        recs = _cells(rng, 77)
        n = 80  # padded to NSSE with zero records, as nln/mac.c:486-520 does
        f = oracle.to_soa(recs[:, :sz], 8)
        flat = btab.reshape(-1)
        g.walk_init_sink(flat, STRIDE)
        q, _ = g.qalloc()
        g.grav_mn(routine, 17, 250, f, n, sz, 0.0, q)
        q2, _ = g.qalloc()
        g.grav_mn(routine, 200, 100, f, n, sz, 0.0, q2)  # overlaps 200..266
        g.walk_terminate_sink(flat, STRIDE)
        ref = np.zeros((300, 4))
        o = oracle.pmn(order, btab[:, 1:4], f, n, sz, quad_phi_sign=qsign)
        ref[17:267] += o[17:267]
        ref[200:300] += o[200:300]
        got = btab[:, 7:11] - before
        _close(got[17:300], ref[17:300])
        assert np.all(got[:17] == 0)
    finally:
        pgrav.set_quad_phi_sign(1)


def test_grav_mn_pM_sK1(gpu_lib):
    pgrav = gpu_lib
    rng = np.random.default_rng(5)
    g = pgrav.GravCuda()
    btab = _bodies(rng, 128, scale=1.0)
    src = np.zeros((64, 4), np.float32)
    src[:, 0] = rng.uniform(0.1, 1.0, 64)
    src[:, 1:4] = btab[rng.integers(0, 128, 64), 1:4] + rng.normal(size=(64, 3)) * 0.2
    eps = 0.3
    f = oracle.to_soa(src, 8)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mn("pM_sK1", 0, 128, f, 64, 4, 1.0 / eps, q)
    g.walk_terminate_sink(flat, STRIDE)
    ref = oracle.pm_sk1(btab[:, 1:4], f, 64, 1.0 / eps)
    _close(btab[:, 7:11], ref)


def test_grav_mns_pMM_sK1(gpu_lib):
    pgrav = gpu_lib
    rng = np.random.default_rng(6)
    g = pgrav.GravCuda()
    btab = _bodies(rng, 500, scale=0.5)
    seg = np.array([[10, 40], [200, 7], [300, 120], [499, 1]], np.uint32)
    eps = 0.05
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mns("pMM_sK1", 64, 130, seg.reshape(-1), int(seg[:, 1].sum()), 0.002, 1.0 / eps, q)
    g.walk_terminate_sink(flat, STRIDE)
    ref = oracle.pmm_sk1(btab[64:194, 1:4], seg.reshape(-1), btab[:, 1:4], 0.002, 1.0 / eps)
    _close(btab[64:194, 7:11], ref)


def test_grav_mnss_queue_layout(gpu_lib):
    """Batched p-p with the queue layout of nln/mac.c:352-392: groups of sinks,
    a zero-work gap group, per-group segment runs and source counts."""
    pgrav = gpu_lib
    rng = np.random.default_rng(8)
    g = pgrav.GravCuda()
    N = 2000
    btab = _bodies(rng, N, scale=0.5)
    groups = [(100, 37), (137, 64), (230, 20), (250, 300)]  # gap 201..229
    ss_index, ss_seg, seg, source_n = [], [], [], []
    m = 0
    base = groups[0][0]
    for first, cnt in groups:
        if first > base + m:
            gap = first - (base + m)
            ss_seg.append((len(seg), 0))
            source_n.append(0)
            ss_index += [len(ss_seg) - 1] * gap
            m += gap
        segs = [(int(b), int(l)) for b, l in zip(rng.integers(0, N - 64, 5), rng.integers(1, 64, 5))]
        ss_seg.append((len(seg), len(segs)))
        source_n.append(sum(l for _, l in segs))
        seg += segs
        ss_index += [len(ss_seg) - 1] * cnt
        m += cnt
    eps = 0.02
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mnss("pMM_mnss_sK1", base, m, ss_index, np.array(ss_seg).reshape(-1),
                np.array(seg).reshape(-1), 0, source_n, 0.0005, 1.0 / eps, q)
    g.walk_terminate_sink(flat, STRIDE)
    ref, bad = oracle.pmm1_mnss_sk1(btab[base:base + m, 1:4], ss_index,
                                    np.array(ss_seg).reshape(-1), np.array(seg).reshape(-1),
                                    btab[:, 1:4], source_n, 0.0005, 1.0 / eps)
    assert bad == 0
    got = btab[base:base + m, 7:11]
    live = np.abs(ref).max(axis=1) > 0
    _close(got[live], ref[live])
    assert np.all(got[~live] == 0)


def test_queue_kernels_split_long_lists(gpu_lib):
    """Few sinks with long source lists: the queue kernels cut every sink's list
    over many waves (up to 128 slices here) and must still equal the oracle."""
    pgrav = gpu_lib
    rng = np.random.default_rng(11)
    g = pgrav.GravCuda()
    N = 6000
    btab = _bodies(rng, N, scale=0.5)
    # pMM_mnss_sK1: 70 sinks in two runs of ~5000 bodies, segments of 1..56 bodies
    ss_seg, seg, source_n = [], [], []
    for _ in range(2):
        segs, tot = [], 0
        while tot < 5000:
            b, ln = int(rng.integers(0, N - 56)), int(rng.integers(1, 57))
            segs.append((b, ln))
            tot += ln
        ss_seg.append((len(seg), len(segs)))
        source_n.append(tot)
        seg += segs
    ss_index = [0] * 33 + [1] * 37
    eps = 0.02
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mnss("pMM_mnss_sK1", 500, 70, ss_index, np.array(ss_seg).reshape(-1),
                np.array(seg).reshape(-1), 0, source_n, 0.0005, 1.0 / eps, q)
    g.walk_terminate_sink(flat, STRIDE)
    ref, bad = oracle.pmm1_mnss_sk1(btab[500:570, 1:4], ss_index, np.array(ss_seg).reshape(-1),
                                    np.array(seg).reshape(-1), btab[:, 1:4], source_n, 0.0005,
                                    1.0 / eps)
    assert bad == 0
    _close(btab[500:570, 7:11], ref)
    # pQ1_qnss: 90 sinks over runs of 2000 and 1200 order-2 records
    btab2 = _bodies(rng, 200)
    recs = _cells(rng, 3200)
    buf = oracle.to_soa(recs[:, :13], 8)
    sg = np.zeros((90, 2), np.uint32)
    sg[:50] = (0, 2000)
    sg[50:] = (2000, 1200)
    flat2 = btab2.reshape(-1)
    g.walk_init_sink(flat2, STRIDE)
    q, _ = g.qalloc()
    g.grav_qnss("pQ1_qnss", 10, 90, sg.reshape(-1), buf, 3200, q)
    g.walk_terminate_sink(flat2, STRIDE)
    ref2 = oracle.pq1_mnss(btab2[10:100, 1:4], sg.reshape(-1), buf)
    _close(btab2[10:100, 7:11], ref2)


def test_grav_mnss_count_mismatch_is_reported(gpu_lib):
    pgrav = gpu_lib
    g = pgrav.GravCuda()
    rng = np.random.default_rng(9)
    btab = _bodies(rng, 64)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mnss("pMM_mnss_sK1", 0, 64, np.zeros(64), np.array([0, 1]), np.array([0, 10]), 0,
                [11], 0.001, 10.0, q)  # source_n says 11, the segments hold 10
    with pytest.raises(pgrav.PgravError, match="source_n"):
        g.walk_terminate_sink(flat, STRIDE)


def test_grav_qnss(gpu_lib):
    pgrav = gpu_lib
    rng = np.random.default_rng(10)
    g = pgrav.GravCuda()
    btab = _bodies(rng, 200)
    recs = _cells(rng, 48)
    buf = oracle.to_soa(recs[:, :13], 8)  # packed SoA-8 order-2 buffer
    seg = np.zeros((120, 2), np.uint32)
    seg[:60] = (0, 24)
    seg[60:100] = (24, 24)
    seg[100:] = (8, 40)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_qnss("pQ1_qnss", 40, 120, seg.reshape(-1), buf, 48, q)
    g.walk_terminate_sink(flat, STRIDE)
    ref = oracle.pq1_mnss(btab[40:160, 1:4], seg.reshape(-1), buf)
    _close(btab[40:160, 7:11], ref)


def _hexacell_table(recs):
    """nln/physics.h:140-157 byte layout (160 B per record)."""
    dt = np.dtype([("mass", "<f4"), ("pos", "<f4", 3), ("rcrit_m", "<f4"), ("bmax", "<f4"),
                   ("lvl_dau", "<i8"), ("bptr", "<u8"), ("rcrit_q", "<f4"), ("rcrit_h", "<f4"),
                   ("q", "<f4", 3), ("q2", "<f4", 5), ("q3", "<f4", 7), ("q4", "<f4", 9),
                   ("u", "<f4", 4)])
    assert dt.itemsize == 160
    t = np.zeros(len(recs), dt)
    t["mass"] = recs[:, 0]
    t["pos"] = recs[:, 1:4]
    t["bmax"] = recs[:, 4]
    t["q"] = recs[:, 5:8]
    t["q2"] = recs[:, 8:13]
    t["q3"] = recs[:, 13:20]
    t["q4"] = recs[:, 20:29]
    return t


def test_grav_qns_with_walk_init_src(gpu_lib):
    pgrav = gpu_lib
    rng = np.random.default_rng(11)
    g = pgrav.GravCuda()
    recs = _cells(rng, 50)
    tab = _hexacell_table(recs)
    raw = np.frombuffer(tab.tobytes(), np.uint8).copy()
    g.walk_init_src(raw, 160, 50)
    btab = _bodies(rng, 90)
    idx = rng.permutation(50)[:33].astype(np.int32)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_qns("pQ1", 5, 80, idx, q)
    g.walk_terminate_sink(flat, STRIDE)
    g.walk_terminate_src()
    ref = oracle.pqq1(btab[5:85, 1:4], idx, recs[:, :13].astype(np.float32))
    _close(btab[5:85, 7:11], ref)


def test_qalloc_slots_and_unknown_routine(gpu_lib):
    pgrav = gpu_lib
    g = pgrav.GravCuda()
    assert g.vecwidth() == 4
    q, inuse = g.qalloc()
    assert 0 <= q < 32 and 0 <= inuse < 32
    rng = np.random.default_rng(12)
    btab = _bodies(rng, 16)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    g.grav_mn("pX", 0, 16, np.zeros(32, np.float32), 8, 4, 0.0, q)  # ignored, as nln/pgrav.cu
    g.walk_terminate_sink(flat, STRIDE)
    assert np.all(btab[:, 7:11] == 0)
    with pytest.raises(pgrav.PgravError):
        g.grav_qnss("pQ2_qnss", 0, 1, np.array([0, 0]), np.zeros(13 * 8, np.float32), 8, q)


def test_py_pgrav_monopole_known_answer(gpu_lib):
    """nln/py_pgrav/test_pgrav.py:6-34 inputs (np.random.seed(123), one sink at
    the origin, 16 sources of mass 1/16); the FP64 brute force pins the answer."""
    pgrav = gpu_lib
    import pgrav_cu
    rs = np.random.RandomState(123)
    n = 16
    source = np.empty((n, 4), np.float32)
    source[:, 0] = 1.0 / n
    source[:, 1] = rs.uniform(0.0, 1.0, n)
    source[:, 2] = rs.uniform(-0.5, 0.5, n)
    source[:, 3] = rs.uniform(-0.5, 0.5, n)
    sink = np.array([1.0, 0.0, 0.0, 0.0], np.float32)
    d = source[:, 1:4].astype(np.float64)  # source - sink
    r = np.linalg.norm(d, axis=1)
    m = source[:, 0].astype(np.float64)
    exact = np.r_[(m[:, None] * d / r[:, None] ** 3).sum(0), -(m / r).sum()]
    accp = np.zeros(4, np.float32)
    pgrav.monopole(sink, accp, 4, source.flatten())
    np.testing.assert_allclose(accp, exact, rtol=0, atol=2e-6 * np.abs(exact).max())
    swiz = source.reshape(-1, 8, 4).transpose(0, 2, 1).flatten()  # test_pgrav.py:30
    accp2 = np.zeros(4, np.float32)
    pgrav.vec_monopole(sink, accp2, 4, np.ascontiguousarray(swiz))
    np.testing.assert_allclose(accp2, exact, rtol=0, atol=2e-6 * np.abs(exact).max())
    pgrav_cu.init()
    accp3 = np.zeros(4, np.float32)
    pgrav_cu.monopole(sink, accp3, 4, np.ascontiguousarray(swiz))
    np.testing.assert_allclose(accp3, exact, rtol=0, atol=2e-6 * np.abs(exact).max())


def _many_calls(pgrav, rng, btab, calls):
    """Runs (routine, first, m, f, n, sz, e) calls on rotating slots; returns the
    results added into the body table."""
    g = pgrav.GravCuda()
    flat = btab.reshape(-1)
    before = btab[:, 7:11].copy()
    g.walk_init_sink(flat, STRIDE)
    for routine, first, m, f, n, sz, e in calls:
        q = -1
        while q < 0:
            q, _ = g.qalloc(count=False)
        if routine == "pMM_sK1":
            g.grav_mns(routine, first, m, f, n, 0.001, e, q)
        else:
            g.grav_mn(routine, first, m, f, n, sz, e, q)
    g.walk_terminate_sink(flat, STRIDE)
    out = btab[:, 7:11] - before
    btab[:, 7:11] = before
    return out.astype(np.float64)


def test_deferred_batches_fill_and_flush(gpu_lib, deferral):
    """Lists over several 16 MB batches, a call larger than a batch (which runs at
    once between the deferred ones), every kind mixed and sinks shared between
    calls: the batched results equal one launch per call up to the FP32 order of
    the sums, and an oracle spot check pins both."""
    if deferral == 0:
        pytest.skip("compares the two modes; runs once")
    pgrav = gpu_lib
    rng = np.random.default_rng(21)
    nb = 4096
    btab = _bodies(rng, nb, scale=2.0)
    recs = _cells(rng, 64)
    pools = {sz: oracle.to_soa(np.tile(recs[:, :sz], (40, 1)), 8) for sz in (4, 13, 29)}
    calls = []
    for k in range(150):
        kind = ("pH", "pQ", "pM", "pM_sK1", "pMM_sK1")[k % 5]
        first = int(rng.integers(0, nb - 100))
        m = int(rng.integers(1, 100))
        if kind == "pMM_sK1":
            seg = np.array([[rng.integers(0, nb - 600), rng.integers(1, 600)],
                            [rng.integers(0, nb - 50), rng.integers(0, 50)]], np.uint32)
            calls.append((kind, first, m, seg.reshape(-1), int(seg[:, 1].sum()), 0, 4.0))
            continue
        sz = {"pH": 29, "pQ": 13, "pM": 4, "pM_sK1": 4}[kind]
        n = 8 * int(rng.integers(1, 320))
        calls.append((kind, first, m, pools[sz], n, sz, 2.0 if kind == "pM_sK1" else 0.0))
    big = oracle.to_soa(np.tile(recs[:, :29], (2600, 1)), 8)  # 166,400 sources: 19 MB
    calls.insert(70, ("pH", 5, 40, big, big.size // 29, 29, 0.0))
    got = _many_calls(pgrav, rng, btab, calls)
    old = pgrav.set_dropin_deferral(0)
    try:
        ref = _many_calls(pgrav, rng, btab, calls)
    finally:
        pgrav.set_dropin_deferral(old)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-30
    assert np.abs(got - ref).max() / scale.max() < 1e-5
    assert (np.abs(got - ref) / scale).max() < 1e-4
    # oracle spot check: the first pH call alone
    kind, first, m, f, n, sz, e = calls[0]
    one = _many_calls(pgrav, rng, btab, [calls[0]])
    o = oracle.pmn(4, btab[:, 1:4], f, n, sz)
    _close(one[first:first + m], o[first:first + m])


def test_deferred_calls_flush_before_the_source_table_changes(gpu_lib, deferral):
    """pQQ1 calls queued against one WalkInitSrcCUDA table are evaluated before the
    next table replaces it."""
    pgrav = gpu_lib
    rng = np.random.default_rng(22)
    g = pgrav.GravCuda()
    btab = _bodies(rng, 90)
    flat = btab.reshape(-1)
    recs_a, recs_b = _cells(rng, 40), _cells(rng, 40)
    idx = rng.permutation(40)[:25].astype(np.int32)
    g.walk_init_sink(flat, STRIDE)
    for recs in (recs_a, recs_b):
        raw = np.frombuffer(_hexacell_table(recs).tobytes(), np.uint8).copy()
        g.walk_init_src(raw, 160, 40)
        q, _ = g.qalloc()
        g.grav_qns("pQ1", 3, 80, idx, q)
    g.walk_terminate_sink(flat, STRIDE)
    g.walk_terminate_src()
    ref = sum(oracle.pqq1(btab[3:83, 1:4], idx, r[:, :13].astype(np.float32))
              for r in (recs_a, recs_b))
    _close(btab[3:83, 7:11], ref)


def test_call_too_large_to_defer_runs_at_once(gpu_lib):
    """A call of more than 65,536 tasks (here 313 sink blocks x 235 source chunks)
    takes the one-launch path even when deferral is on; sampled sinks against the
    oracle."""
    pgrav = gpu_lib
    rng = np.random.default_rng(23)
    g = pgrav.GravCuda()
    m, n = 20000, 120000
    btab = _bodies(rng, m, scale=1.0)
    src = np.zeros((n, 4), np.float32)
    src[:, 0] = rng.uniform(0.5, 1.5, n) / n
    src[:, 1:4] = rng.uniform(-3, 3, (n, 3))
    f = oracle.to_soa(src, 8)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, STRIDE)
    q, _ = g.qalloc()
    g.grav_mn("pM", 0, m, f, n, 4, 0.0, q)
    g.walk_terminate_sink(flat, STRIDE)
    pick = rng.choice(m, 300, replace=False)
    ref = oracle.pmn(0, btab[pick, 1:4], f, n, 4)
    _close(btab[pick, 7:11], ref, tol=1e-4)
