"""torch.autograd.Function over the batched evaluation and its adjoint kernel.

    plan = batch.Plan(pos0, mass0, theta=0.6, order=4, eps=1e-4)   # tree and lists fixed here
    acc = gravity(pos, mass, plan)        # (N, 4) float32: ax, ay, az, phi
    loss = (acc[:, :3] * w).sum() + (acc[:, 3] * v).sum()
    loss.backward()                       # pos.grad, mass.grad from the adjoint kernels

Forward moves the plan's bodies to `pos`/`mass` (same tree topology, groups and
lists; moments recomputed in FP64) and runs the FP32 group-relative compensated
kernel.  Backward runs the adjoint with the incoming gradient as the weights
(w_i, v_i).  GPU tensors stay on the device (pgrav_plan_*_device, ordered on
torch's current stream); CPU tensors go through host arrays.  The plan is
shared and mutable: forward saves its inputs and the plan's version, and a
backward that finds the plan holding other bodies (a later forward or update)
puts this forward's bodies back first.
"""
import torch

from .batch import REL32C


def _on_gpu(t):
    return t.is_cuda


class _Stream:
    """torch's current stream for the library's device entry points.  torch's
    default stream is the legacy null stream, which the library's own
    (non-blocking) stream does not wait for: with it, the device calls are
    bracketed by synchronisations instead."""

    def __init__(self, plan):
        self.plan = plan
        self.handle = torch.cuda.current_stream().cuda_stream

    def __enter__(self):
        if not self.handle:
            torch.cuda.current_stream().synchronize()
        return self.handle or None

    def __exit__(self, *exc):
        if not self.handle:
            self.plan.synchronize()
        return False


def _update(plan, pos, mass):
    if _on_gpu(pos) and _on_gpu(mass):
        p = pos.detach().to(torch.float64).contiguous()
        m = mass.detach().to(torch.float64).contiguous()
        with _Stream(plan) as st:
            plan.update_device(p.data_ptr(), m.data_ptr(), st)
    else:
        plan.update(pos.detach().cpu().double().numpy(), mass.detach().cpu().double().numpy())


class TreeGravity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass, plan):
        _update(plan, pos, mass)
        ctx.plan = plan
        ctx.version = plan.version
        ctx.save_for_backward(pos, mass)
        if _on_gpu(pos):
            acc = torch.empty((plan.n, 4), dtype=torch.float32, device=pos.device)
            with _Stream(plan) as st:
                plan.evaluate(REL32C, stream=st)
                plan.result_device(REL32C, acc.data_ptr(), fp64=False, stream=st)
            return acc
        plan.evaluate(REL32C)
        return torch.from_numpy(plan.result(REL32C)).to(dtype=torch.float32)

    @staticmethod
    def backward(ctx, grad):
        pos, mass = ctx.saved_tensors
        plan = ctx.plan
        if plan.version != ctx.version:
            _update(plan, pos, mass)
            ctx.version = plan.version
        if _on_gpu(grad):
            w = grad.detach().to(torch.float32).contiguous()
            g = torch.empty((plan.n, 4), dtype=torch.float64, device=grad.device)
            with _Stream(plan) as st:
                plan.adjoint_device(w.data_ptr(), g.data_ptr(), w_fp64=False, stream=st)
        else:
            g = torch.from_numpy(plan.adjoint(grad.detach().cpu().double().numpy()))
        gpos = g[:, :3].to(device=pos.device, dtype=pos.dtype)
        gmass = g[:, 3].to(device=mass.device, dtype=mass.dtype)
        return gpos, gmass, None


def gravity(pos, mass, plan):
    return TreeGravity.apply(pos, mass, plan)
