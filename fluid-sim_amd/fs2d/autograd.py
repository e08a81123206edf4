"""torch.autograd wrapper of a checkpointed rollout (fs2d_adjoint_* of the C
ABI): a thin harness, so that a PyTorch optimisation loop can treat n steps of
the all-fluid semi-Lagrangian simulation as a differentiable function of its
initial state (u, v, density).

The rollout and its reverse pass run on the device through the C ABI; the
tensors cross the boundary through host memory (the fields are the
simulation's own device buffers, behind the handle).
"""
import numpy as np
import torch

from . import F_U, F_V, F_DENSITY


class Rollout(torch.autograd.Function):
    """(u0, v0, d0) -> (u_n, v_n, d_n) after `nsteps` steps of adj.sim."""

    @staticmethod
    def forward(ctx, u0, v0, d0, adj, nsteps):
        sim = adj.sim
        sim.set(F_U, u0.detach().cpu().numpy())
        sim.set(F_V, v0.detach().cpu().numpy())
        sim.set(F_DENSITY, d0.detach().cpu().numpy())
        adj.forward(nsteps)
        ctx.adj = adj
        dev = u0.device
        return tuple(torch.from_numpy(sim.get(f)).to(dev) for f in (F_U, F_V, F_DENSITY))

    @staticmethod
    def backward(ctx, gu, gv, gd):
        adj = ctx.adj
        sim = adj.sim
        for f, g in ((F_U, gu), (F_V, gv), (F_DENSITY, gd)):
            if g is None:
                adj.set(f, np.zeros(sim.shape(f), dtype=np.float32))
            else:
                adj.set(f, g.detach().float().cpu().numpy())
        adj.backward()
        dev = gu.device if gu is not None else gd.device
        grads = tuple(torch.from_numpy(adj.get(f)).to(dev) for f in (F_U, F_V, F_DENSITY))
        return grads + (None, None)


def rollout(adj, u0, v0, d0, nsteps):
    return Rollout.apply(u0, v0, d0, adj, nsteps)
