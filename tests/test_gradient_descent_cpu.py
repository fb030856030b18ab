"""The gradient method's descent extension (umans_policy::gradient_iterations, not in the
reference) in the oracle, pinned by mathematics:

* force-model costs have the quadratic cost 0.5 |v - target|^2 / dt
  (ForceBasedFunction.cpp:26-36), so the descent v <- v - eta dt (v - target) / dt has
  the closed form v_N = target + (1 - eta)^N (v - target); with GoalReachingForce alone,
  target = v + (pref - v) / max(relax, dt) / mass * dt (GoalReachingForce.cpp:26-29);
* one iteration of step 1 is the reference's single gradient step up to rounding;
* eta = 0 leaves the velocity unchanged, and the iteration count 1 is the reference's
  path bit for bit.
"""
import numpy as np
import pytest

import oracle_py as O
from umans_amd import scenarios as S
from umans_amd.registry import pack_policy


def lone_agents(n=40, seed=3):
    """Agents 100 m apart (no neighbours in range, no contact), random velocities."""
    r = np.random.default_rng(seed)
    a = S._agents(n)
    a["pos"][:, 0] = (100.0 * np.arange(n)).astype(np.float32)
    a["goal"][:] = (a["pos"] + r.uniform(-30, 30, (n, 2))).astype(np.float32)
    a["vel"][:] = r.uniform(-1.0, 1.0, (n, 2)).astype(np.float32)
    a["pref_speed"][:] = r.uniform(1.0, 1.4, n).astype(np.float32)
    a["max_speed"][:] = 1e4
    a["max_accel"][:] = 1e4
    a["viewdir"] = S.initial_viewdir(a["pos"], a["goal"])
    return a


def goal_policy(iters, eta, relax=0.5, mass=1.0):
    return pack_policy("gradient", [("GoalReachingForce", 1.0, {"agentMass": mass})], relaxation_time=relax,
                       gradient_iterations=iters, gradient_step=eta)


@pytest.mark.parametrize("iters,eta", [(2, 0.5), (20, 0.1), (7, 1.3), (50, 0.05)])
def test_descent_matches_closed_form_fp64(iters, eta):
    dt, relax, mass = 0.1, 0.5, 1.7
    a = lone_agents()
    w = O.make_world(0, 0, 0, dt, 0)
    o = O.step(w, [goal_policy(iters, eta, relax, mass)], a, precision=1)
    v = a["vel"].astype(np.float64)
    dtf = float(np.float32(dt))
    d = a["goal"].astype(np.float64) - a["pos"].astype(np.float64)
    pref = d / np.linalg.norm(d, axis=1, keepdims=True) * a["pref_speed"].astype(np.float64)[:, None]
    target = v + (pref - v) / max(float(np.float32(relax)), dtf) / float(np.float32(mass)) * dtf
    vN = target + (1.0 - float(np.float32(eta))) ** iters * (v - target)
    # acceleration (v_N - v) / dt, clamps far away: next velocity = v_N; no contact
    np.testing.assert_allclose(o["vel64"], vN, rtol=0, atol=1e-12)


def test_one_iteration_of_step_one_is_the_reference_step():
    dt = 0.1
    a = lone_agents()
    w = O.make_world(0, 0, 0, dt, 0)
    ref = O.step(w, [goal_policy(1, 1.0)], a, precision=1)          # the reference's path
    # descent with two iterations of step 0.5 on a quadratic cost: v_2 = target + 0.25 (v - target),
    # while the reference's step lands on the target (eta 1): the extension changes results
    two = O.step(w, [goal_policy(2, 0.5)], a, precision=1)
    assert not np.allclose(two["vel64"], ref["vel64"], atol=1e-6)
    # eta 0 leaves every velocity where it was
    zero = O.step(w, [goal_policy(5, 0.0)], a, precision=1)
    np.testing.assert_array_equal(zero["vel64"], a["vel"].astype(np.float64))
    # iterations 0 and 1 are the reference's path bit for bit, in FP32 too
    for prec, key in ((0, "vel"), (1, "vel64")):
        r1 = O.step(w, [goal_policy(1, 0.3)], a, precision=prec)
        r0 = O.step(w, [goal_policy(0, 0.3)], a, precision=prec)
        r = O.step(w, [pack_policy("gradient", [("GoalReachingForce", 1.0, {"agentMass": 1.0})],
                                   relaxation_time=0.5)], a, precision=prec)
        np.testing.assert_array_equal(r1[key], r[key])
        np.testing.assert_array_equal(r0[key], r[key])


def test_descent_in_fp32_follows_fp64_on_a_crowd():
    """Force models with neighbours (SocialForces, PowerLaw): the FP32 oracle's descent
    stays within 1e-5 relative of the FP64 one."""
    pols = [dict(S.socialforces_gradient_policy(), gradient_iterations=20, gradient_step=0.1),
            dict(S.powerlaw_gradient_policy(), gradient_iterations=20, gradient_step=0.1)]
    sc = S.random_cloud(600, 30.0, pols, world_type=1, seed=12)
    w = O.make_world(1, sc.width, sc.height, 0.05, 0)
    o = O.step(w, pols, sc.agents, precision=1)
    o32 = O.step(w, pols, sc.agents, precision=0)
    d = np.linalg.norm(o32["vel"].astype(np.float64) - o["vel64"], axis=1)
    rel = d / np.maximum(np.linalg.norm(o["vel64"], axis=1), 1e-3)
    assert np.quantile(rel, 0.99) < 1e-5, np.sort(rel)[-5:]
