"""CPU checks of the adjoint's FP64 reference (tests/adjoint_ref.py), for both
interaction models it differentiates (SocialForcesAvoidance, PowerLaw):

* its forward step equals the FP64 oracle's step (which is pinned to the compiled
  reference);
* its autograd gradient agrees with central differences of its own forward rollout,
  taken with steps small enough to stay on one smooth piece of the loss;
* its float32 branch decisions (the FP32 kernels' decisions) equal its float64 ones on
  states with no value at a threshold, and a planted mistake in one is caught;
* neighbour lists read back from a record give the step the brute-force lists give;
* its reverse sweep along a recorded trajectory equals autograd through the whole
  rollout.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as AR
import oracle_py as O
from umans_amd import scenarios as S

MODELS = ("socialforces", "powerlaw")
# differentiated parameters at the start point (A = V0, B = sigma; k, tau0) and the
# target-generating ones
START = {"socialforces": (1500.0, 0.12), "powerlaw": (2.0, 2.5)}
TRUE = {"socialforces": (2000.0, 0.08), "powerlaw": (1.5, 3.0)}
KIND = {"socialforces": 2, "powerlaw": 1}


def small(model, n=300, size=17.3, seed=9):
    p0, p1 = START[model]
    if model == "powerlaw":
        return S.calibration(n=n, size=size, seed=seed, model="powerlaw", k=p0, tau0=p1)
    return S.calibration(n=n, size=size, seed=seed, V0=p0, sigma=p1)


def params_of(sc, model):
    s0, s1 = S.PARAM_SLOTS[KIND[model]]
    pol = sc.policies[0]
    c = [i for i in range(pol["n_cost"]) if pol["cost_kind"][i] == KIND[model]][0]
    return (torch.tensor(float(np.float32(pol["cost_params"][c][s0])), dtype=AR.F64),
            torch.tensor(float(np.float32(pol["cost_params"][c][s1])), dtype=AR.F64))


def tensors(sc):
    return (torch.tensor(sc.agents["pos"], dtype=AR.F64), torch.tensor(sc.agents["vel"], dtype=AR.F64))


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("seed", [9, 10])
def test_fp64_reference_step_matches_oracle(model, seed):
    sc = small(model, seed=seed)
    c = AR.consts_from(sc)
    p0, p1 = params_of(sc, model)
    pos, vel = tensors(sc)
    p_1, v_1 = AR.step(pos, vel, p0, p1, c)
    o = O.step(O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors), sc.policies,
               sc.agents, precision=1)
    np.testing.assert_allclose(v_1.numpy(), o["vel64"], rtol=0, atol=1e-9)
    if model == "powerlaw":
        # the PowerLaw crowd must exercise the model: many pairs inside the time horizon
        D = AR.f32_decisions(sc.agents["pos"], sc.agents["vel"], *[x.numpy() for x in
                             AR.neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])], c,
                             float(p1))
        assert int(D["act"].sum()) > sc.n // 2


@pytest.mark.parametrize("model", MODELS)
def test_fp64_reference_gradient_matches_differences(model):
    sc = small(model, n=120, size=11.0)
    c = AR.consts_from(sc)
    pos0, vel0 = tensors(sc)
    q0, q1 = TRUE[model]
    with torch.no_grad():
        tg, pos, vel = [], pos0, vel0
        for _ in range(8):
            pos, vel = AR.step(pos, vel, torch.tensor(q0, dtype=AR.F64), torch.tensor(q1, dtype=AR.F64), c)
            tg.append(pos)
        targets = torch.stack(tg)
    a0, a1 = START[model]
    P0 = torch.tensor(a0, dtype=AR.F64, requires_grad=True)
    P1 = torch.tensor(a1, dtype=AR.F64, requires_grad=True)
    v0 = vel0.clone().requires_grad_(True)
    L = AR.rollout_loss(pos0, v0, P0, P1, c, targets)
    g0, g1, gv = torch.autograd.grad(L, [P0, P1, v0])
    assert abs(float(g0)) > 0 and abs(float(g1)) > 0

    def f(x0, x1, v):
        with torch.no_grad():
            return float(AR.rollout_loss(pos0, v, torch.tensor(x0, dtype=AR.F64), torch.tensor(x1, dtype=AR.F64),
                                         c, targets))
    # the loss is piecewise smooth (view-factor switches, cut-offs, list changes): a step
    # of 1e-9 relative stays on one smooth piece
    for name, g, h in (("p0", g0, 1e-9 * a0), ("p1", g1, 1e-9 * a1)):
        dx = (h, 0.0) if name == "p0" else (0.0, h)
        fd = (f(a0 + dx[0], a1 + dx[1], vel0) - f(a0 - dx[0], a1 - dx[1], vel0)) / (2 * h)
        assert abs(fd - float(g)) <= 1e-5 * abs(fd) + 1e-12, (model, name, fd, float(g))
    # a directional derivative in the initial velocities: a step of 1e-8 (at 1e-9 the
    # rounding of the loss, ~1e-16 L / 2h, is already 1e-5 of the SocialForces
    # derivative; at 1e-7 the PowerLaw rollout crosses one of its cut-offs)
    w = torch.from_numpy(np.random.default_rng(1).standard_normal(vel0.shape))
    h = 1e-8
    fd = (f(a0, a1, vel0 + h * w) - f(a0, a1, vel0 - h * w)) / (2 * h)
    assert abs(fd - float((gv * w).sum())) <= 1e-5 * abs(fd) + 1e-12, (model, fd, float((gv * w).sum()))


def decisions_both(model, seed=11):
    # the calibration crowd scattered uniformly, so that agents overlap (contact decisions)
    sc = small(model, n=400, size=20.0, seed=seed)
    sc.agents["pos"] = np.random.default_rng(seed).uniform(-10.0, 10.0, (sc.n, 2)).astype(np.float32)
    c = AR.consts_from(sc)
    p0, p1 = params_of(sc, model)
    pos, vel = tensors(sc)
    idx, off, mask = AR.neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])
    d32 = AR.f32_decisions(sc.agents["pos"], sc.agents["vel"], idx.numpy(), off.numpy(), mask.numpy(), c, float(p1))
    d64 = AR.f64_decisions(pos, vel, idx, off, mask, c, float(p1))
    return d32, d64, mask


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("seed", [11, 12])
def test_f32_branch_decisions_equal_f64_ones(model, seed):
    """On crowds with no value at a threshold the float32 decisions (the FP32 kernels'
    expressions) and the float64 ones agree mask for mask, and every kind of decision
    goes both ways somewhere in the crowd."""
    d32, d64, mask = decisions_both(model, seed)
    for k in ("reached", "inrange", "act", "contact"):
        assert torch.equal(d32[k], d64[k]), k
    # the view factor also differs where the float cosine falls outside [-1, 1] (a
    # neighbour straight behind or ahead): the FP32 path counts it as in view
    vd = (d32["view"] != d64["view"]) & d64["act"]
    assert torch.equal(vd & ~d32["view_oob"], torch.zeros_like(vd))
    listed = mask.sum()
    assert 0 < int(d64["act"].sum()) < int(listed)
    assert 0 < int(d64["contact"].sum()) < int(listed)
    if model == "socialforces":
        assert 0 < int((d64["view"] & d64["act"]).sum()) < int(d64["act"].sum())


def test_f32_branch_decisions_catch_a_planted_mistake(monkeypatch):
    """A view factor decided on the wrong side of 100 degrees: the comparison above sees
    it."""
    monkeypatch.setattr(AR, "VIEW_COS", -AR.VIEW_COS)
    d32, _, _ = decisions_both("socialforces")
    monkeypatch.undo()
    _, d64, _ = decisions_both("socialforces")
    assert not torch.equal(d32["view"] & d32["act"], d64["view"] & d64["act"])


def test_lists_from_record_equal_bruteforce_lists():
    sc = small("socialforces", n=300, size=16.0, seed=12)
    c = AR.consts_from(sc)
    p0, p1 = params_of(sc, "socialforces")
    pos, vel = tensors(sc)
    o = O.step(O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors), sc.policies, sc.agents,
               precision=1, list_stride=128)
    rec = AR.lists_from_record(o["nb_count"], o["nb_id"], o["nb_seg"], c["W"], c["H"])
    bf = AR.neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])
    assert (o["nb_seg"][o["nb_id"] >= 0] > 0).any()            # the crowd crosses the seams
    for x, y in zip(AR.step(pos, vel, p0, p1, c, lists=rec), AR.step(pos, vel, p0, p1, c, lists=bf)):
        np.testing.assert_allclose(x.numpy(), y.numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("model", MODELS)
def test_replay_gradient_equals_autograd_through_the_rollout(model):
    """The reverse sweep's bookkeeping (loss gradients of each step's output, parameter
    sums, step order) along a float64 trajectory equals autograd through the rollout."""
    sc = small(model, n=150, size=12.0, seed=13)
    c = AR.consts_from(sc)
    pos0, vel0 = tensors(sc)
    q0, q1 = TRUE[model]
    a0, a1 = START[model]
    T = 6
    with torch.no_grad():
        tg, pos, vel = [], pos0, vel0
        for _ in range(T):
            pos, vel = AR.step(pos, vel, torch.tensor(q0, dtype=AR.F64), torch.tensor(q1, dtype=AR.F64), c)
            tg.append(pos)
        targets = torch.stack(tg)
        states, lists, pos, vel = [(pos0.numpy(), vel0.numpy())], [], pos0, vel0
        for _ in range(T):
            lst = AR.neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])
            lists.append(lst)
            pos, vel = AR.step(pos, vel, torch.tensor(a0, dtype=AR.F64), torch.tensor(a1, dtype=AR.F64), c,
                               lists=lst)
            states.append((pos.numpy(), vel.numpy()))
    loss, g0, g1, gp, gv = AR.replay_gradient(states, lists, targets.numpy(), a0, a1, c, branches="f64")
    P0 = torch.tensor(a0, dtype=AR.F64, requires_grad=True)
    P1 = torch.tensor(a1, dtype=AR.F64, requires_grad=True)
    p0r = pos0.clone().requires_grad_(True)
    v0r = vel0.clone().requires_grad_(True)
    L = AR.rollout_loss(p0r, v0r, P0, P1, c, targets)
    r0, r1, rp, rv = torch.autograd.grad(L, [P0, P1, p0r, v0r])
    assert loss == pytest.approx(float(L), rel=1e-12)
    assert g0 == pytest.approx(float(r0), rel=1e-9)
    assert g1 == pytest.approx(float(r1), rel=1e-9)
    np.testing.assert_allclose(gv, rv.numpy(), rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(gp, rp.numpy(), rtol=1e-9, atol=1e-14)


def test_given_clamp_branches():
    """Clamp flags equal to the float64 decisions give the float64 step; a flipped speed
    clamp changes it."""
    sc = small("socialforces", n=300, size=17.3, seed=9)
    c = AR.consts_from(sc)
    p0, p1 = params_of(sc, "socialforces")
    pos, vel = tensors(sc)
    lists = AR.neighbor_lists(pos, c["world_type"], c["W"], c["H"], c["rng"], c["cap"])
    seen = []
    orig = AR._clamp

    def rec(v, L, clamped=None):
        seen.append((AR._norm(v) > L).numpy())
        return orig(v, L, clamped)
    AR._clamp = rec
    try:
        ref = AR.step(pos, vel, p0, p1, c, lists=lists)
    finally:
        AR._clamp = orig
    flags = seen[0].astype(np.uint8) | (seen[1].astype(np.uint8) << 1)
    assert flags.any() and not (flags == 3).all()
    got = AR.step(pos, vel, p0, p1, c, lists=lists, clamps=flags)
    for x, y in zip(ref, got):
        np.testing.assert_array_equal(x.numpy(), y.numpy())
    flip = flags ^ 2
    got = AR.step(pos, vel, p0, p1, c, lists=lists, clamps=flip)
    assert np.abs(got[1].numpy() - ref[1].numpy()).max() > 1e-6
