"""The hand-written adjoint of the smooth gradient step on the GPU.

* against the FP64 reference gradient (tests/adjoint_ref.py: float64 torch autograd of
  the same step) on a small calibration crowd: loss, dL/dA, dL/dB and dL/dv0;
* checkpoint interval: the gradients do not depend on it;
* the torch.autograd.Function wrapper.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as AR
from umans_amd import adjoint as ADJ, gpu, scenarios as S

pytestmark = pytest.mark.gpu

T_SMALL = 20


def small():
    return S.calibration(n=324, size=18.0, seed=9)


def gpu_world(sc):
    return gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)


def gpu_grad(sc, targets, A, B, T, ck):
    w = gpu_world(sc)
    w.set_policies([S.with_sf_params(p, A, B) for p in sc.policies])
    w.upload(sc.agents)
    loss = w.adjoint_forward(targets, T, ck)
    dp, dpos, dvel = w.adjoint_backward(1)
    w.close()
    return loss, dp[0], dvel


def targets_for(sc, T):
    w = gpu_world(sc)
    tg = ADJ.rollout_targets(w, sc, 2000.0, 0.08, T)
    w.close()
    return tg


@pytest.mark.parametrize("T", [1, 5, T_SMALL])
def test_adjoint_matches_fp64_reference(T):
    """Over horizons on which the FP32 and FP64 rollouts of this crowd stay within
    about 1e-4 m of each other."""
    sc = small()
    tg = targets_for(sc, T)
    loss, dAB, dv = gpu_grad(sc, tg, 1500.0, 0.12, T, 10)
    c = AR.consts_from(sc)
    p0 = torch.tensor(sc.agents["pos"], dtype=AR.F64)
    v0 = torch.tensor(sc.agents["vel"], dtype=AR.F64, requires_grad=True)
    A = torch.tensor(1500.0, dtype=AR.F64, requires_grad=True)
    B = torch.tensor(np.float32(0.12), dtype=AR.F64, requires_grad=True)
    L = AR.rollout_loss(p0, v0, A, B, c, torch.tensor(tg, dtype=AR.F64))
    gA, gB, gv = torch.autograd.grad(L, [A, B, v0])
    assert abs(loss - float(L.detach())) <= 1e-4 * abs(float(L.detach()))
    assert abs(dAB[0] - float(gA)) <= 1e-3 * abs(float(gA)), (dAB[0], float(gA))
    assert abs(dAB[1] - float(gB)) <= 1e-3 * abs(float(gB)), (dAB[1], float(gB))
    gv = gv.numpy()
    assert np.linalg.norm(dv - gv) <= 1e-3 * np.linalg.norm(gv), np.linalg.norm(dv - gv) / np.linalg.norm(gv)


def test_checkpoint_interval_does_not_change_gradients():
    sc = small()
    tg = targets_for(sc, T_SMALL)
    ref = gpu_grad(sc, tg, 1500.0, 0.12, T_SMALL, 10)
    for ck in (1, 7, T_SMALL):
        g = gpu_grad(sc, tg, 1500.0, 0.12, T_SMALL, ck)
        assert g[0] == pytest.approx(ref[0], rel=1e-12)       # the forward is deterministic
        # the reverse pass sums neighbour contributions with atomics: rounding order only
        np.testing.assert_allclose(g[1], ref[1], rtol=1e-4)
        assert np.linalg.norm(g[2] - ref[2]) <= 1e-4 * np.linalg.norm(ref[2])


def test_autograd_function():
    sc = small()
    tg = targets_for(sc, T_SMALL)
    w = gpu_world(sc)
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    v0 = torch.tensor(sc.agents["vel"], device=dev, requires_grad=True)
    params = torch.tensor([[1500.0, 0.12]], device=dev, requires_grad=True)
    L = ADJ.RolloutLoss.apply(v0, params, w, sc, tg, T_SMALL, 10)
    L.backward()
    loss, dAB, dv = gpu_grad(sc, tg, 1500.0, 0.12, T_SMALL, 10)
    assert float(L) == pytest.approx(loss, rel=1e-12)
    np.testing.assert_allclose(params.grad.cpu().numpy()[0], dAB, rtol=1e-4)
    assert np.linalg.norm(v0.grad.cpu().numpy() - dv) <= 1e-4 * np.linalg.norm(dv)
    w.close()
