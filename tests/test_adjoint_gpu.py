"""Adjoint of the batched evaluation (tree and lists fixed) on the GPU.

1. Central finite differences (h = 1e-5, FP64 forward through the same lists
   with moments recomputed for the moved body) on 64 sampled coordinates and
   masses agree with the FP32 adjoint.
2. At N = 8192 the tree adjoint approximates the oracle's FP64 direct-sum
   gradient to the tree's own accuracy.
3. The torch.autograd.Function returns the same gradients, from CPU tensors
   through host arrays and from GPU tensors through the device entry points.
4. Two calls give bit-identical gradients (no atomics).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _loss(plan, w4, variant="f64"):
    plan.evaluate(variant)
    return float((plan.result(variant) * w4).sum())


@pytest.mark.parametrize("mode", [1, 5])
def test_adjoint_matches_finite_differences(gpu_lib, mode):
    """mode 5: bisection groups, whose sinks the adjoint kernels read and write
    through the slot table."""
    from pgrav import batch, ics
    pos, mass = ics.clustered(6000, seed=31, nhalo=8)
    rng = np.random.default_rng(5)
    w4 = rng.normal(size=(pos.shape[0], 4))
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-3, soft="plummer", merge_groups=mode,
                      rcb_cell=1024)
    grad = plan.adjoint(w4)
    h = 1e-5
    picks = rng.choice(pos.shape[0], 16, replace=False)
    fd, ad = [], []
    for k in picks:
        for q in range(4):
            p1, p2 = pos.copy(), pos.copy()
            m1, m2 = mass.copy(), mass.copy()
            if q < 3:
                p1[k, q] += h
                p2[k, q] -= h
                step = h
            else:
                m1[k] *= 1 + h
                m2[k] *= 1 - h
                step = mass[k] * h
            plan.update(p1, m1)
            lp = _loss(plan, w4)
            plan.update(p2, m2)
            lm = _loss(plan, w4)
            fd.append((lp - lm) / (2 * step))
            ad.append(grad[k, q])
    plan.update(pos, mass)
    fd, ad = np.array(fd).reshape(-1, 4), np.array(ad).reshape(-1, 4)
    for q, name in ((slice(0, 3), "position"), (slice(3, 4), "mass")):
        scale = np.sqrt((fd[:, q] ** 2).mean())
        err = np.abs(ad[:, q] - fd[:, q]).max() / scale
        assert err < 2e-3, (name, err)
    plan.close()


def test_adjoint_approximates_direct_sum_gradient(gpu_lib):
    from pgrav import batch, ics
    pos, mass = ics.clustered(8192, seed=32, nhalo=8)
    rng = np.random.default_rng(6)
    w4 = rng.normal(size=(pos.shape[0], 4))
    eps = 2e-3
    plan = batch.Plan(pos, mass, theta=0.4, order=4, eps=eps, soft="plummer")
    grad = plan.adjoint(w4)
    ref = oracle.direct_grad(pos, mass, w4, soft_type=1, eps=eps)
    dx = np.linalg.norm(grad[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    dm = np.abs(grad[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
    assert np.median(dx) < 1e-3, np.median(dx)
    assert np.median(dm) < 1e-3, np.median(dm)
    plan.close()


def test_autograd_function(gpu_lib):
    import torch
    from pgrav import batch, ics
    from pgrav.autograd import gravity
    pos, mass = ics.clustered(4000, seed=33, nhalo=4)
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-3)
    w4 = np.random.default_rng(7).normal(size=(pos.shape[0], 4))
    tp = torch.tensor(pos, dtype=torch.float64, requires_grad=True)
    tm = torch.tensor(mass, dtype=torch.float64, requires_grad=True)
    acc = gravity(tp, tm, plan)
    loss = (acc.double() * torch.tensor(w4)).sum()
    loss.backward()
    ref = plan.adjoint(w4)
    # the same weights, the same fixed-order sums
    np.testing.assert_array_equal(tp.grad.numpy(), ref[:, :3])
    np.testing.assert_array_equal(tm.grad.numpy(), ref[:, 3])
    plan.close()


def test_adjoint_is_deterministic(gpu_lib):
    """Fixed-order sums, no atomics: two adjoint calls give identical gradients,
    also for K1 and unsoftened bodies and bisection groups."""
    from pgrav import batch, ics
    pos, mass = ics.clustered(12000, seed=34, nhalo=6)
    w4 = np.random.default_rng(8).normal(size=(pos.shape[0], 4))
    for soft, eps, mode in (("plummer", 1e-3, 1), ("k1", 2e-3, 5), ("none", 0.0, 5)):
        print("adjoint determinism soft=%s mode=%d" % (soft, mode), flush=True)
        plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=eps, soft=soft, merge_groups=mode,
                          rcb_cell=2048)
        g1 = plan.adjoint(w4)
        g2 = plan.adjoint(w4)
        np.testing.assert_array_equal(g1, g2)
        assert np.isfinite(g1).all()
        plan.close()


def test_autograd_device_path_and_stale_plan(gpu_lib):
    """GPU tensors through the device entry points (no host copies) give the
    host path's gradients; a backward after a second forward on the same plan
    (other positions) still differentiates at its own forward's bodies."""
    import torch
    from pgrav import batch, ics
    from pgrav.autograd import gravity
    pos, mass = ics.clustered(5000, seed=35, nhalo=4)
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-3)
    rng = np.random.default_rng(9)
    w4 = rng.normal(size=(pos.shape[0], 4))
    pos2 = pos + rng.normal(size=pos.shape) * 1e-4
    ref1 = plan.adjoint(w4)  # the plan holds pos
    plan.update(pos2, mass)
    ref2 = plan.adjoint(w4)
    plan.evaluate("rel32c")
    acc2_ref = plan.result("rel32c")
    dev = torch.device("cuda")
    tw = torch.tensor(w4, dtype=torch.float32, device=dev)
    tp1 = torch.tensor(pos, dtype=torch.float64, device=dev, requires_grad=True)
    tm1 = torch.tensor(mass, dtype=torch.float64, device=dev, requires_grad=True)
    tp2 = torch.tensor(pos2, dtype=torch.float64, device=dev, requires_grad=True)
    tm2 = torch.tensor(mass, dtype=torch.float64, device=dev, requires_grad=True)
    acc1 = gravity(tp1, tm1, plan)
    acc2 = gravity(tp2, tm2, plan)  # the plan now holds pos2
    np.testing.assert_array_equal(acc2.detach().cpu().numpy(), acc2_ref.astype(np.float32))
    (acc1 * tw).sum().backward()  # must move the plan back to pos
    (acc2 * tw).sum().backward()
    w32 = w4.astype(np.float32).astype(np.float64)  # the device path's weights are float32
    np.testing.assert_allclose(tp1.grad.cpu().numpy(), plan_adjoint_at(plan, pos, mass, w32)[:, :3],
                               rtol=0, atol=0)
    np.testing.assert_allclose(tp2.grad.cpu().numpy(), plan_adjoint_at(plan, pos2, mass, w32)[:, :3],
                               rtol=0, atol=0)
    scale = np.abs(ref1[:, :3]).max()
    np.testing.assert_allclose(tp1.grad.cpu().numpy(), ref1[:, :3], rtol=0, atol=1e-5 * scale)
    np.testing.assert_allclose(tp2.grad.cpu().numpy(), ref2[:, :3], rtol=0, atol=1e-5 * scale)
    plan.close()


def plan_adjoint_at(plan, pos, mass, w4):
    plan.update(pos, mass)
    return plan.adjoint(w4)
