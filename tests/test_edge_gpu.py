"""Edge cases on the GPU: a single body, sizes straddling the 64-sink group,
many coincident bodies, empty drop-in calls, and rejected updates.  The batched
results are compared with the library's FP64 direct sum (itself equal to the
oracle's, test_batch_gpu.py)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_single_body(gpu_lib):
    from pgrav import batch
    plan = batch.Plan(np.array([[0.3, 0.4, 0.5]]), np.array([2.0]), theta=0.6, order=4,
                      eps=1e-2, G=1.5)
    for v in ("rel32c", "f64"):
        plan.evaluate(v)
        a = plan.result(v)
        ref = plan.direct([0])
        # Plummer self term: no force, potential -G m / eps
        assert np.abs(a[0, :3]).max() == 0.0
        assert abs(a[0, 3] - ref[0, 3]) <= 1e-6 * abs(ref[0, 3])
        assert abs(ref[0, 3] - (-1.5 * 2.0 / 1e-2)) <= 1e-12 * 300
    plan.close()


@pytest.mark.parametrize("n", [2, 63, 64, 65, 129, 1000])
def test_small_and_ragged_sizes(gpu_lib, n):
    from pgrav import batch
    rng = np.random.default_rng(n)
    pos = rng.uniform(size=(n, 3))
    mass = rng.uniform(0.5, 1.5, n) / n
    plan = batch.Plan(pos, mass, theta=0.5, order=4, eps=1e-3)
    ref = plan.direct(np.arange(n))
    for v, tol in (("rel32c", 1e-3), ("f64", 1e-3)):
        plan.evaluate(v)
        a = plan.result(v)
        da = np.linalg.norm(a[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
        assert da.max() < tol, (v, da.max())
    plan.close()


def test_coincident_bodies(gpu_lib):
    from pgrav import batch
    rng = np.random.default_rng(7)
    pos = np.concatenate([np.full((300, 3), 0.25), rng.uniform(size=(40, 3))])
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    plan = batch.Plan(pos, mass, theta=0.5, order=4, eps=1e-3)
    ref = plan.direct(np.arange(pos.shape[0]))
    plan.evaluate("rel32c")
    a = plan.result("rel32c")
    scale = np.linalg.norm(ref[:, :3], axis=1).max()
    assert np.abs(a[:, :3] - ref[:, :3]).max() < 1e-4 * scale
    assert np.all(np.isfinite(a))
    plan.close()


def test_update_rejects_bad_bodies(gpu_lib):
    from pgrav import batch
    from pgrav._lib import PgravError
    rng = np.random.default_rng(2)
    pos, mass = rng.uniform(size=(500, 3)), np.full(500, 0.002)
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-3)
    bad = mass.copy()
    bad[10] = 0.0
    with pytest.raises(PgravError):
        plan.update(pos, bad)
    p2 = pos.copy()
    p2[4, 0] = np.nan
    with pytest.raises(PgravError):
        plan.update(p2, mass)
    plan.update(pos, mass)  # still usable
    plan.evaluate("rel32c")
    assert np.all(np.isfinite(plan.result("rel32c")))
    plan.close()


def test_dropin_empty_calls(gpu_lib):
    pgrav = gpu_lib
    rng = np.random.default_rng(3)
    g = pgrav.GravCuda()
    btab = np.zeros((40, 14), np.float32)
    btab[:, 0] = 1.0 / 40
    btab[:, 1:4] = rng.uniform(-1, 1, (40, 3))
    before = btab.copy()
    f = oracle.to_soa(np.zeros((8, 29), np.float32), 8)
    flat = btab.reshape(-1)
    g.walk_init_sink(flat, 14)
    q, _ = g.qalloc()
    g.grav_mn("pH", 0, 0, f, 8, 29, 0.0, q)        # no sinks
    q2, _ = g.qalloc()
    g.grav_mn("pM", 0, 40, f[:0], 0, 4, 0.0, q2)   # no sources
    g.walk_terminate_sink(flat, 14)
    np.testing.assert_array_equal(btab, before)
