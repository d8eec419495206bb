"""The reference's dispatch on the batched path's own lists.

For sampled sink groups, the group's order-4 cells go through the drop-in
grav_mn_CUDA("pH") as SoA-8 Hvec blocks and its bodies through
grav_mns_CUDA("pMM_sK1") as MMvec segments -- the calls nln/mac.c:1218 and the
leaf p-p path make -- and the result added into the body table must equal the
batched kernel's (K1 softening for bodies, unsoftened cells, as pgrav.cu).
Both paths are FP32; the drop-in path uses absolute coordinates like the
reference, so it is compared with the batched FP64 result at the precision
absolute FP32 coordinates allow.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dropin_dispatch_reproduces_batched_result(gpu_lib):
    pgrav = gpu_lib
    from pgrav import batch, ics, reflists
    pos, mass = ics.clustered(20000, seed=41, nhalo=12)
    eps = 2e-3
    kw = dict(theta=0.6, order=4, eps=eps, soft="k1")
    H = batch.HostPlan(pos, mass, **kw)
    perm, spos = H.sorted()
    plan = batch.Plan(pos, mass, **kw)
    plan.evaluate("f64")
    ref = plan.result("f64")[perm]
    plan.evaluate("abs32")
    abs32 = plan.result("abs32")[perm]
    plan.close()

    n = pos.shape[0]
    btab = np.zeros((n, 14), np.float32)
    btab[:, 0] = mass[perm]
    btab[:, 1:4] = spos
    flat = btab.reshape(-1)
    g = pgrav.GravCuda()
    g.walk_init_sink(flat, 14)
    groups = np.random.default_rng(4).choice(H.stats()["ngroups"], 40, replace=False)
    rows = []
    for gi in groups:
        first, cnt, _ = H.group(int(gi))
        q, _ = g.qalloc()
        soa, source_n = reflists.group_soa(H, int(gi), "hexa")
        if source_n:
            g.grav_mn("pH", first, cnt, soa, source_n, 29, 0.0, q)
        seg, nb = reflists.group_segments(H, int(gi))
        if nb:
            g.grav_mns("pMM_sK1", first, cnt, seg.reshape(-1), nb, float(mass[0]), 1.0 / eps, q)
        rows.append(np.arange(first, first + cnt))
    g.walk_terminate_sink(flat, 14)
    rows = np.concatenate(rows)
    got = btab[rows, 7:11].astype(np.float64)
    da_ref = np.linalg.norm(got[:, :3] - ref[rows, :3], axis=1) / np.linalg.norm(ref[rows, :3], axis=1)
    da_abs = np.linalg.norm(got[:, :3] - abs32[rows, :3], axis=1) / np.linalg.norm(abs32[rows, :3], axis=1)
    assert np.median(da_ref) < 1e-5 and da_ref.max() < 1e-3, (np.median(da_ref), da_ref.max())
    assert np.median(da_abs) < 1e-5, np.median(da_abs)
    dp = np.abs(got[:, 3] - ref[rows, 3]) / np.abs(ref[rows, 3])
    assert dp.max() < 1e-4, dp.max()
