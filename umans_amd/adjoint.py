"""Differentiable rollouts of the smooth gradient step as a torch.autograd.Function.

`RolloutLoss.apply(v0, params, world, scenario, targets, nsteps, checkpoint_every,
param_kind)` returns the position loss of an nsteps rollout (mean over steps and agents
of the squared minimum-image distance to `targets`) and backpropagates to the initial
velocities `v0` (n x 2) and to the two parameters of the interaction model of every
policy (`params`, n_policies x 2): (A, B) = (V0, sigma) of SocialForcesAvoidance
(param_kind 2) or (k, tau0) of PowerLaw (param_kind 1). The forward pass runs on the
GPU through the C ABI and keeps a checkpoint every `checkpoint_every` steps; the
backward pass recomputes each segment and runs the hand-written reverse kernels.

Device tensors stay on the device: the targets are read and dL/dv0 is written by the
library at their device addresses. `params` and `v0` (2 floats per policy, n x 2)
cross to the host, because the policy table and the packed agent state are built
there (umans_set_policies, umans_upload).
"""
import numpy as np
import torch

from .scenarios import with_params


def _as_f32(t):
    return t.detach().to(torch.float32).contiguous()


class RolloutLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, params, world, scenario, targets, nsteps, checkpoint_every=10, param_kind=2):
        p = params.detach().to("cpu", torch.float64).reshape(-1, 2).numpy()
        pols = [with_params(pk, param_kind, p[i, 0], p[i, 1]) for i, pk in enumerate(scenario.policies)]
        world.set_policies(pols)
        agents = dict(scenario.agents)
        agents["vel"] = _as_f32(v0).cpu().numpy()
        world.upload(agents)
        if torch.is_tensor(targets) and targets.is_cuda:
            tg = _as_f32(targets)
            if tuple(tg.shape) != (nsteps, world.n, 2):
                raise ValueError(f"targets must have shape {(nsteps, world.n, 2)}")
            torch.cuda.current_stream(tg.device).synchronize()     # the library reads them on its own stream
            ctx.keep = tg
            loss = world.adjoint_forward_ptr(tg.data_ptr(), nsteps, checkpoint_every, param_kind)
        else:
            tg = targets.detach().cpu().numpy() if torch.is_tensor(targets) else targets
            loss = world.adjoint_forward(np.asarray(tg, np.float32), nsteps, checkpoint_every, param_kind)
        ctx.world = world
        ctx.n_pol = len(pols)
        ctx.dev = v0.device
        return torch.tensor(loss, dtype=torch.float64, device=v0.device)

    @staticmethod
    def backward(ctx, grad):
        g = float(grad)
        if ctx.dev.type == "cuda":
            dvel = torch.empty((ctx.world.n, 2), dtype=torch.float32, device=ctx.dev)
            torch.cuda.current_stream(ctx.dev).synchronize()
            dp = ctx.world.adjoint_backward_ptr(ctx.n_pol, 0, dvel.data_ptr())   # synchronises its stream
            dv = dvel * g
        else:
            dp, _dpos, dvel = ctx.world.adjoint_backward(ctx.n_pol)
            dv = torch.from_numpy(dvel) * g
        dpar = torch.from_numpy(dp).to(ctx.dev) * g
        return dv, dpar, None, None, None, None, None, None


def rollout_targets(world, scenario, p0, p1, nsteps, param_kind=2):
    """Positions after each of nsteps steps (nsteps x n x 2, upload order) with the
    interaction parameters (p0, p1) ((V0, sigma) or (k, tau0), by param_kind)."""
    world.set_policies([with_params(p, param_kind, p0, p1) for p in scenario.policies])
    world.upload(scenario.agents)
    out = np.zeros((nsteps, scenario.n, 2), np.float32)
    for t in range(nsteps):
        world.step(1)
        out[t] = world.download()["pos"]
    return out


def record_rollout(world, scenario, p0, p1, nsteps, param_kind=2, stride=128):
    """The FP32 trajectory and neighbour lists of an nsteps rollout from the scenario's
    agents with parameters (p0, p1): states[t] = (pos, vel) before step t (t = 0..nsteps,
    upload order) and lists[t] = (count, ids, segments) of step t (umans_dump_neighbors)."""
    world.set_policies([with_params(p, param_kind, p0, p1) for p in scenario.policies])
    world.upload(scenario.agents)
    world.record_neighbors(True)
    a = world.download()
    states = [(a["pos"], a["vel"])]
    lists = []
    for _ in range(nsteps):
        world.step(1)
        cnt, ids, _d2, seg = world.neighbors(stride=stride)
        K = max(int(cnt.max()), 1)
        lists.append((cnt, ids[:, :K].copy(), seg[:, :K].astype(np.int8)))
        a = world.download()
        states.append((a["pos"], a["vel"]))
    world.record_neighbors(False)
    return states, lists
