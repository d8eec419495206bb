"""FP64 upward pass on the GPU (pgrav_plan_update): after the bodies move under
a fixed tree, every node's centre of mass, raw moments and bmax come from the
device kernels.  Checked per node against the oracle's moments summed directly
from the node's bodies at that centre (the parallel-axis identity pins the
multinomial shift: a wrong coefficient or index fails it), and the forces
after the update against a plan built from scratch on the moved bodies where
the trees coincide (no body crosses a cell boundary)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_update_moments_match_direct_moments(gpu_lib):
    from pgrav import batch, ics
    pos, mass = ics.clustered(30000, seed=41, nhalo=16)
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4)
    H = batch.HostPlan(pos, mass, theta=0.6, order=4, eps=1e-4)  # same tree, host side
    rng = np.random.default_rng(3)
    pos2 = pos + rng.normal(scale=2e-3, size=pos.shape)
    mass2 = mass * rng.uniform(0.5, 1.5, size=mass.shape)
    plan.update(pos2, mass2)
    perm = plan.permutation()
    sp, sm = pos2[perm], mass2[perm]
    nn = plan.stats()["nnodes"]
    c4, bm, raw = plan.cells()
    for nd in np.concatenate([[0], rng.choice(nn, 300, replace=False)]):
        c = H.cell(int(nd))
        b = sp[c["first"]:c["first"] + c["count"]]
        m = sm[c["first"]:c["first"] + c["count"]]
        cm = (m[:, None] * b).sum(0) / m.sum()
        np.testing.assert_allclose(c4[nd, :3], cm, rtol=0, atol=1e-13)
        assert abs(c4[nd, 3] - m.sum()) <= 1e-13 * m.sum()
        ref = oracle.raw_moments(b, m, c4[nd, :3])
        scale = np.abs(ref).max()
        assert np.abs(raw[nd] - ref).max() <= 1e-11 * scale, nd
        # bmax is an upper bound on the bodies' distance from the centre
        r = np.sqrt(((b - c4[nd, :3]) ** 2).sum(1)).max() if c["count"] > 1 else 0.0
        assert bm[nd] >= r * (1 - 1e-12)
    plan.close()


def test_update_forces_match_a_fresh_plan(gpu_lib):
    """Bodies displaced well inside their leaves: the fresh build has the same
    tree, groups and lists, so the two results agree to FP32 round-off."""
    from pgrav import batch, ics
    pos, mass = ics.clustered(20000, seed=42, nhalo=8)
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4)
    rng = np.random.default_rng(4)
    pos2 = pos + rng.normal(scale=1e-9, size=pos.shape)
    plan.update(pos2, mass)
    plan.evaluate("rel32c")
    a = plan.result("rel32c")
    fresh = batch.Plan(pos2, mass, theta=0.6, order=4, eps=1e-4)
    assert fresh.stats()["pairs_total"] == plan.stats()["pairs_total"]
    fresh.evaluate("rel32c")
    b = fresh.result("rel32c")
    rel = np.linalg.norm(a[:, :3] - b[:, :3], axis=1) / np.linalg.norm(b[:, :3], axis=1)
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4, (np.median(rel), rel.max())
    plan.close()
    fresh.close()
