"""CPU checks of the adjoint's FP64 reference (tests/adjoint_ref.py): its forward step
equals the FP64 oracle's step (which is pinned to the compiled reference), and its
autograd gradient agrees with central differences of its own forward rollout taken
with steps small enough to stay on one smooth piece of the loss."""
import numpy as np
import pytest
import torch

import adjoint_ref as AR
import oracle_py as O
from umans_amd import scenarios as S


def small(n=300, size=17.3, seed=9):
    return S.calibration(n=n, size=size, seed=seed)


@pytest.mark.parametrize("seed", [9, 10])
def test_fp64_reference_step_matches_oracle(seed):
    sc = small(seed=seed)
    c = AR.consts_from(sc)
    pol = sc.policies[0]
    A = torch.tensor(pol["cost_params"][0][2], dtype=AR.F64)
    B = torch.tensor(np.float32(pol["cost_params"][0][3]), dtype=AR.F64)
    p0 = torch.tensor(sc.agents["pos"], dtype=AR.F64)
    v0 = torch.tensor(sc.agents["vel"], dtype=AR.F64)
    p1, v1 = AR.step(p0, v0, A, B, c)
    o = O.step(O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors), sc.policies,
               sc.agents, precision=1)
    np.testing.assert_allclose(v1.numpy(), o["vel64"], rtol=0, atol=1e-9)


def test_fp64_reference_gradient_matches_differences():
    sc = small(n=120, size=11.0)
    c = AR.consts_from(sc)
    p0 = torch.tensor(sc.agents["pos"], dtype=AR.F64)
    v0 = torch.tensor(sc.agents["vel"], dtype=AR.F64)
    with torch.no_grad():
        tg, pos, vel = [], p0, v0
        for _ in range(8):
            pos, vel = AR.step(pos, vel, torch.tensor(2000.0, dtype=AR.F64), torch.tensor(0.08, dtype=AR.F64), c)
            tg.append(pos)
        targets = torch.stack(tg)
    A = torch.tensor(1500.0, dtype=AR.F64, requires_grad=True)
    B = torch.tensor(0.12, dtype=AR.F64, requires_grad=True)
    L = AR.rollout_loss(p0, v0, A, B, c, targets)
    gA, gB = torch.autograd.grad(L, [A, B])
    # the loss is piecewise smooth (view-factor switches, cut-offs, list changes): a step
    # of 1e-9 relative stays on one smooth piece
    for name, g, x0, h in (("A", gA, 1500.0, 1.5e-6), ("B", gB, 0.12, 1.2e-10)):
        def f(x):
            a = torch.tensor(x if name == "A" else 1500.0, dtype=AR.F64)
            b = torch.tensor(x if name == "B" else 0.12, dtype=AR.F64)
            with torch.no_grad():
                return float(AR.rollout_loss(p0, v0, a, b, c, targets))
        fd = (f(x0 + h) - f(x0 - h)) / (2 * h)
        assert abs(fd - float(g)) <= 1e-5 * abs(fd) + 1e-12, (name, fd, float(g))
