"""Differentiable rollouts of the smooth gradient step as a torch.autograd.Function.

`RolloutLoss.apply(v0, params, world, scenario, targets, nsteps, checkpoint_every)`
returns the position loss of an nsteps rollout (mean over steps and agents of the
squared minimum-image distance to `targets`) and backpropagates to the initial
velocities `v0` (n x 2) and to the SocialForces parameters `params` = (A, B) = (V0,
sigma) of every policy (n_policies x 2). The forward pass runs on the GPU through the
C ABI and keeps a checkpoint every `checkpoint_every` steps; the backward pass
recomputes each segment and runs the hand-written reverse kernels. torch only carries
tensors across the boundary.
"""
import numpy as np
import torch

from . import gpu
from .scenarios import with_sf_params


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, params, world, scenario, targets, nsteps, checkpoint_every=10):
        p = params.detach().cpu().numpy().astype(np.float64).reshape(-1, 2)
        pols = [with_sf_params(pk, p[i, 0], p[i, 1]) for i, pk in enumerate(scenario.policies)]
        world.set_policies(pols)
        agents = dict(scenario.agents)
        agents["vel"] = v0.detach().cpu().numpy().astype(np.float32)
        world.upload(agents)
        loss = world.adjoint_forward(np.asarray(targets, np.float32), nsteps, checkpoint_every, param_kind=2)
        ctx.world = world
        ctx.n_pol = len(pols)
        ctx.dev = v0.device
        return torch.tensor(loss, dtype=torch.float64, device=v0.device)

    @staticmethod
    def backward(ctx, grad):
        dp, _dpos, dvel = ctx.world.adjoint_backward(ctx.n_pol)
        g = float(grad)
        dv = torch.from_numpy(dvel).to(ctx.dev) * g
        dpar = torch.from_numpy(dp).to(ctx.dev) * g
        return dv, dpar, None, None, None, None, None


def rollout_targets(world, scenario, V0, sigma, nsteps):
    """Positions after each of nsteps steps (nsteps x n x 2, upload order) with the
    SocialForces parameters (V0, sigma)."""
    world.set_policies([with_sf_params(p, V0, sigma) for p in scenario.policies])
    world.upload(scenario.agents)
    out = np.zeros((nsteps, scenario.n, 2), np.float32)
    for t in range(nsteps):
        world.step(1)
        out[t] = world.download()["pos"]
    return out
