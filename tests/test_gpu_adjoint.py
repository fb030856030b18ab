"""The hand-written adjoint of the smooth gradient step on the GPU, for both interaction
models it differentiates: SocialForcesAvoidance (gradients in A = V0, B = sigma) and
PowerLaw (gradients in k, tau0), each with GoalReachingForce, and the initial velocities.

* against the float64 reverse sweep along the GPU's own recorded FP32 trajectory
  (tests/adjoint_ref.py: replay_gradient, the exact derivative of the map the adjoint
  differentiates, with the FP32 step's lists and branch decisions), on a small crowd at
  horizons 1, 5, 20 and on the full calibration crowd (10k agents, 100 steps,
  checkpoints every 10);
* against an independent FP64 rollout (brute-force lists, float64 branches) with torch
  autograd, at horizons on which the FP32 and FP64 rollouts stay together;
* checkpoint interval: the gradients do not depend on it;
* the torch.autograd.Function, with device tensors.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as AR
from umans_amd import adjoint as ADJ, gpu, scenarios as S

pytestmark = pytest.mark.gpu

T_SMALL = 20
KIND = {"socialforces": 2, "powerlaw": 1}
START = {"socialforces": (1500.0, 0.12), "powerlaw": (2.0, 2.5)}
TRUE = {"socialforces": (2000.0, 0.08), "powerlaw": (1.5, 3.0)}
DEV = "cuda" if torch.cuda.is_available() else "cpu"
START_F32 = tuple(float(np.float32(x)) for x in (1500.0, 0.12))


def small(model):
    if model == "powerlaw":
        return S.calibration(n=324, size=18.0, seed=9, model="powerlaw", k=START[model][0], tau0=START[model][1])
    return S.calibration(n=324, size=18.0, seed=9)


def gpu_world(sc):
    return gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)


def gpu_grad(sc, model, targets, T, ck, params=None):
    p0, p1 = params or START[model]
    w = gpu_world(sc)
    try:
        w.set_policies([S.with_params(p, KIND[model], p0, p1) for p in sc.policies])
        w.upload(sc.agents)
        loss = w.adjoint_forward(targets, T, ck, param_kind=KIND[model])
        dp, dpos, dvel = w.adjoint_backward(1)
        clamps = w.adjoint_clamps(T)
    finally:
        w.close()
    return loss, dp[0], dpos, dvel, clamps


def targets_for(sc, model, T):
    w = gpu_world(sc)
    try:
        return ADJ.rollout_targets(w, sc, *TRUE[model], T, param_kind=KIND[model])
    finally:
        w.close()


def replay(sc, model, targets, T, clamps):
    """FP64 reverse sweep along the GPU's FP32 trajectory at the start parameters, with
    the clamp branches its reverse pass reported."""
    w = gpu_world(sc)
    try:
        states, lists = ADJ.record_rollout(w, sc, *START[model], T, param_kind=KIND[model])
    finally:
        w.close()
    c = AR.consts_from(sc)
    rec = [AR.lists_from_record(cnt, ids, seg, c["W"], c["H"]) for cnt, ids, seg in lists]
    p0, p1 = (float(np.float32(x)) for x in START[model])
    return AR.replay_gradient(states, rec, targets, p0, p1, c, device=DEV, clamps=clamps), states, rec, c


def rel(a, b):
    return abs(a - b) / abs(b)


def l2rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("model", ["socialforces", "powerlaw"])
@pytest.mark.parametrize("T", [1, 5, T_SMALL])
def test_adjoint_matches_fp64_sweep_along_its_trajectory(model, T):
    sc = small(model)
    tg = targets_for(sc, model, T)
    loss, dP, dpos, dv, clamps = gpu_grad(sc, model, tg, T, 10)
    (rl, g0, g1, gp, gv), _, _, _ = replay(sc, model, tg, T, clamps)
    assert abs(g0) > 0 and abs(g1) > 0
    assert rel(loss, rl) <= 1e-5
    assert rel(dP[0], g0) <= 1e-3, (dP[0], g0)
    assert rel(dP[1], g1) <= 1e-3, (dP[1], g1)
    assert l2rel(dv, gv) <= 1e-3, l2rel(dv, gv)
    assert l2rel(dpos, gp) <= 1e-3, l2rel(dpos, gp)


@pytest.mark.parametrize("model,T", [("socialforces", 1), ("socialforces", 5), ("socialforces", T_SMALL),
                                     ("powerlaw", 1), ("powerlaw", 5)])
def test_adjoint_matches_independent_fp64_rollout(model, T):
    """Over horizons on which the FP32 and FP64 rollouts of the crowd stay within about
    1e-4 m of each other: the FP64 rollout has its own brute-force lists and branches."""
    sc = small(model)
    tg = targets_for(sc, model, T)
    loss, dP, _, dv, _ = gpu_grad(sc, model, tg, T, 10)
    c = AR.consts_from(sc)
    p0 = torch.tensor(sc.agents["pos"], dtype=AR.F64)
    v0 = torch.tensor(sc.agents["vel"], dtype=AR.F64, requires_grad=True)
    P0 = torch.tensor(float(np.float32(START[model][0])), dtype=AR.F64, requires_grad=True)
    P1 = torch.tensor(float(np.float32(START[model][1])), dtype=AR.F64, requires_grad=True)
    L = AR.rollout_loss(p0, v0, P0, P1, c, torch.tensor(tg, dtype=AR.F64))
    g0, g1, gv = torch.autograd.grad(L, [P0, P1, v0])
    assert rel(loss, float(L.detach())) <= 1e-4
    assert rel(dP[0], float(g0)) <= 1e-3, (dP[0], float(g0))
    assert rel(dP[1], float(g1)) <= 1e-3, (dP[1], float(g1))
    assert l2rel(dv, gv.numpy()) <= 1e-3, l2rel(dv, gv.numpy())


def test_adjoint_calibration_crowd_full_size():
    """The calibration workload itself: 10k agents in a 100 x 100 m toric world,
    SocialForces + GoalReachingForce, 100 steps, checkpoints every 10; gradients in
    (A, B) from (1500, 0.12) and in the 10k initial velocities, against the float64
    reverse sweep along the GPU's trajectory."""
    sc = S.calibration()
    T = 100
    tg = targets_for(sc, "socialforces", T)
    loss, dP, dpos, dv, clamps = gpu_grad(sc, "socialforces", tg, T, 10)
    (rl, g0, g1, gp, gv), states, rec, c = replay(sc, "socialforces", tg, T, clamps)
    print(f"calibration10000 T=100: loss {loss:.6g} (fp64 {rl:.6g}); dL/dA {dP[0]:.6g} (fp64 {g0:.6g}); "
          f"dL/dB {dP[1]:.6g} (fp64 {g1:.6g}); |dL/dv0| {np.linalg.norm(dv):.6g} (fp64 {np.linalg.norm(gv):.6g}, "
          f"rel {l2rel(dv, gv):.2e})", flush=True)
    # one step from a recorded mid-rollout state: the GPU's reverse step and the float64
    # one agree (the lists and decisions the replay uses are the GPU's)
    t = T // 2
    a1 = dict(sc.agents)
    a1["pos"], a1["vel"] = states[t]
    sc1 = S.Scenario(sc.name, sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors, sc.policies, a1)
    tg1 = (states[t + 1][0] + np.random.default_rng(0).normal(0.0, 0.05, states[t][0].shape)).astype(np.float32)[None]
    l1, dP1, dp1, dv1, cl1 = gpu_grad(sc1, "socialforces", tg1, 1, 1)
    assert np.array_equal(cl1[0], clamps[t])
    rl1, h0, h1, hp, hv = AR.replay_gradient([states[t], states[t + 1]], [rec[t]], tg1, *START_F32, c, device=DEV,
                                             clamps=cl1)
    assert rel(l1, rl1) <= 1e-9 and rel(dP1[1], h1) <= 1e-4
    assert l2rel(dv1, hv) <= 1e-4 and l2rel(dp1, hp) <= 1e-4
    assert rel(loss, rl) <= 1e-5
    assert rel(dP[0], g0) <= 1e-3, (dP[0], g0)
    assert rel(dP[1], g1) <= 1e-3, (dP[1], g1)
    assert l2rel(dv, gv) <= 1e-3, l2rel(dv, gv)


@pytest.mark.parametrize("model", ["socialforces", "powerlaw"])
def test_checkpoint_interval_does_not_change_gradients(model):
    sc = small(model)
    tg = targets_for(sc, model, T_SMALL)
    ref = gpu_grad(sc, model, tg, T_SMALL, 10)
    for ck in (1, 7, T_SMALL):
        g = gpu_grad(sc, model, tg, T_SMALL, ck)
        assert g[0] == pytest.approx(ref[0], rel=1e-12)       # the forward is deterministic
        assert np.array_equal(g[4], ref[4])                    # and so are its clamp branches
        # the reverse pass sums neighbour contributions with atomics: rounding order only
        np.testing.assert_allclose(g[1], ref[1], rtol=1e-4)
        assert np.linalg.norm(g[3] - ref[3]) <= 1e-4 * np.linalg.norm(ref[3])


@pytest.mark.parametrize("model", ["socialforces", "powerlaw"])
def test_autograd_function(model):
    sc = small(model)
    tg = targets_for(sc, model, T_SMALL)
    w = gpu_world(sc)
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    v0 = torch.tensor(sc.agents["vel"], device=dev, requires_grad=True)
    params = torch.tensor([list(START[model])], device=dev, requires_grad=True)
    L = ADJ.RolloutLoss.apply(v0, params, w, sc, torch.tensor(tg, device=dev), T_SMALL, 10, KIND[model])
    L.backward()
    loss, dP, _, dv, _ = gpu_grad(sc, model, tg, T_SMALL, 10)
    assert float(L.detach()) == pytest.approx(loss, rel=1e-12)
    np.testing.assert_allclose(params.grad.cpu().numpy()[0], dP, rtol=1e-4)
    assert np.linalg.norm(v0.grad.cpu().numpy() - dv) <= 1e-4 * np.linalg.norm(dv)
    w.close()
