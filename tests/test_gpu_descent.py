"""The gradient method's descent extension (umans_policy::gradient_iterations) on the
GPU against the oracle: 20 iterations of step 0.1 for force models (the force +
gradient class kernel) and for VanToll (numeric gradient, costFunction.cpp:36-54) and
TtcaDca (analytic gradient, TtcaDca.cpp:131-262) in the generic class, beside a policy
that keeps the reference's single step; infinite and toric worlds, capped and uncapped
lists, one step from the same state. Bar: new velocities within 1e-5 relative of the
FP64 oracle, or of the FP32 oracle where FP32 rounding and FP64 part."""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import REL_TOL, gpu_step, rel_err
from umans_amd import scenarios as S
from umans_amd.registry import pack_policy

pytestmark = pytest.mark.gpu


def descent(p, iters=20, eta=0.1):
    return dict(p, gradient_iterations=iters, gradient_step=eta)


POLICIES = [
    descent(S.socialforces_gradient_policy(rng=4.0)),
    descent(S.powerlaw_gradient_policy(rng=4.0)),
    descent(pack_policy("gradient", [("VanToll", 1.0, {"range": 4.0})])),
    descent(pack_policy("gradient", [("TtcaDca", 1.0, {"range": 4.0})]), iters=8, eta=0.3),
    S.socialforces_gradient_policy(rng=4.0),          # the reference's single step
]


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("cap", [16, 0])
def test_descent_matches_oracle(world_type, cap):
    sc = S.random_cloud(3000, 50.0, POLICIES, world_type=world_type, max_neighbors=cap, seed=90 + cap + world_type,
                        speed=1.2)
    dt = 0.05
    out, _, st = gpu_step(world_type, sc.width, sc.height, dt, POLICIES, sc.agents, max_neighbors=cap,
                          record=False)
    w = O.make_world(world_type, sc.width, sc.height, dt, cap)
    o64 = O.step(w, POLICIES, sc.agents, precision=1)
    o32 = O.step(w, POLICIES, sc.agents, precision=0)
    e64 = rel_err(out["vel"], o64["vel64"])
    e32 = rel_err(out["vel"], o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= REL_TOL)
    pol = sc.agents["policy"]
    for q in range(len(POLICIES)):
        m = pol == q
        print(f"policy {q}: max e64 {e64[m].max():.2e}  max e32 {e32[m].max():.2e}  fails {int((~ok[m]).sum())}")
    assert ok.all(), f"agents {np.where(~ok)[0][:8]} e64 {e64[~ok][:4]} e32 {e32[~ok][:4]}"
    # the descent moved the velocities: not the single step's result
    single = [dict(p, gradient_iterations=1) for p in POLICIES]
    o1 = O.step(w, single, sc.agents, precision=1)
    moved = rel_err(out["vel"], o1["vel64"]) > 1e-3
    assert moved[pol < 4].mean() > 0.5
    assert not moved[pol == 4].any()
