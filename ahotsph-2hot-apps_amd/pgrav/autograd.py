"""torch.autograd.Function over the batched evaluation and its adjoint kernel.

    plan = batch.Plan(pos0, mass0, theta=0.6, order=4, eps=1e-4)   # tree and lists fixed here
    acc = gravity(pos, mass, plan)        # (N, 4) float32: ax, ay, az, phi
    loss = (acc[:, :3] * w).sum() + (acc[:, 3] * v).sum()
    loss.backward()                       # pos.grad, mass.grad from the adjoint kernels

Forward moves the plan's bodies to `pos`/`mass` (same tree topology, groups and
lists; moments recomputed in FP64) and runs the FP32 group-relative compensated
kernel.  Backward runs pgrav_plan_adjoint with the incoming gradient as the
weights (w_i, v_i).  Inputs are float64 tensors (CPU or GPU); the exchange with
the library goes through host arrays.
"""
import torch

from .batch import REL32C


class TreeGravity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass, plan):
        plan.update(pos.detach().cpu().double().numpy(), mass.detach().cpu().double().numpy())
        plan.evaluate(REL32C)
        acc = torch.from_numpy(plan.result(REL32C)).to(device=pos.device, dtype=torch.float32)
        ctx.plan = plan
        ctx.pos_meta = (pos.device, pos.dtype)
        ctx.mass_meta = (mass.device, mass.dtype)
        return acc

    @staticmethod
    def backward(ctx, grad):
        w4 = grad.detach().cpu().double().numpy()
        g = torch.from_numpy(ctx.plan.adjoint(w4))
        gpos = g[:, :3].to(device=ctx.pos_meta[0], dtype=ctx.pos_meta[1])
        gmass = g[:, 3].to(device=ctx.mass_meta[0], dtype=ctx.mass_meta[1])
        return gpos, gmass, None


def gravity(pos, mass, plan):
    return TreeGravity.apply(pos, mass, plan)
