"""The dense bidirectional flow and the batched scenario sweep at their configured sizes,
against the FP32 and FP64 oracles (the parity bar of tests/test_gpu_parity.py):

* bidirectional flow: 100k agents in a 400 x 400 m toric world, Karamouzas with regular
  sampling over 16 angles x 8 speeds plus the base (129 candidates), range 5 m, lists
  capped at 16, dt 0.1 s — neighbour lists bit-exact and new velocities, one step from
  the start and one step from the state the GPU reaches after 10 steps;
* batched sweep: all 512 worlds (20 to 5000 agents each, circle / two-lane / uniform
  layouts, PowerLaw + GoalReachingForce with 12 x 6 sampling, per-world k and tau0,
  uncapped lists, dt 0.1 s) in one batch — every world's new velocities against the
  oracle run on that world alone, one step from the start and one from the state after
  30 batched steps.
"""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import REL_TOL, rel_err
from umans_amd import gpu, scenarios as S

pytestmark = pytest.mark.gpu


def check_step(out_vel, o32, o64, what):
    e64 = rel_err(out_vel, o64["vel64"])
    e32 = rel_err(out_vel, o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    assert ok.all(), f"{what}: {(~ok).sum()} agents off, e64 {np.sort(e64)[-4:]}"


def test_bidirectional_flow_100k():
    sc = S.bidirectional()
    assert sc.n == 100_000 and sc.max_neighbors == 16
    pol = sc.policies[0]
    assert pol["s_angle_samples"] == 16 and pol["s_speed_samples"] == 8 and pol["s_include_base"] == 1
    w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    wd = O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors)
    try:
        a = dict(sc.agents)
        for label, advance in (("start", 0), ("after 10 steps", 10)):
            w.upload(a)
            if advance:
                w.step(advance)
                s = w.download()
                a = dict(sc.agents)
                for k in ("pos", "vel", "viewdir"):
                    a[k] = s[k]
                w.upload(a)
            w.record_neighbors(True)
            w.step(1)
            out = w.download()
            cnt, ids, d2, seg = w.neighbors(stride=16)
            w.record_neighbors(False)
            o32 = O.step(wd, sc.policies, a, precision=0, nthreads=16, list_stride=16)
            o64 = O.step(wd, sc.policies, a, precision=1, nthreads=16)
            np.testing.assert_array_equal(cnt, o32["nb_count"])
            full = cnt > 0
            np.testing.assert_array_equal(np.where(np.arange(16)[None] < cnt[:, None], ids, -1),
                                          np.where(np.arange(16)[None] < cnt[:, None], o32["nb_id"], -1))
            assert full.mean() > 0.99 and (o32["nb_full"] > 16).mean() > 0.5     # most lists hit the cap
            if advance:
                assert np.abs(a["vel"]).max() > 0.5
            check_step(out["vel"], o32, o64, f"bidirectional {label}")
    finally:
        w.close()


def test_batched_sweep_512_worlds():
    sw = S.sweep()
    assert len(sw.worlds) == 512 and sw.sizes.min() >= 20 and sw.sizes.max() <= 5000
    w = gpu.World(0, 0, 0, 0.1, sw.policies, max_neighbors=sw.max_neighbors, batch=sw.worlds)
    try:
        a = dict(sw.agents)
        starts = np.concatenate([[0], np.cumsum(sw.sizes)])
        for label, advance in (("start", 0), ("after 30 steps", 30)):
            w.upload(a)
            if advance:
                w.step(advance)
                s = w.download()
                a = dict(sw.agents)
                for k in ("pos", "vel", "viewdir"):
                    a[k] = s[k]
                w.upload(a)
            w.step(1)
            out = w.download()
            st = w.stats()
            assert st["max_list"] <= 128, st           # lists within the library's capacity here
            for k in range(len(sw.worlds)):
                sl = slice(starts[k], starts[k + 1])
                ak = {key: (v[sl] if v is not None and np.ndim(v) > 0 else v) for key, v in a.items()}
                ak["policy"] = np.zeros(sw.sizes[k], np.int32)
                wt, W, H, dt = sw.worlds[k]
                wd = O.make_world(wt, W, H, dt, sw.max_neighbors)
                o32 = O.step(wd, [sw.policies[k]], ak, precision=0, nthreads=16)
                o64 = O.step(wd, [sw.policies[k]], ak, precision=1, nthreads=16)
                check_step(out["vel"][sl], o32, o64, f"sweep {label} world {k} ({sw.sizes[k]} agents)")
    finally:
        w.close()
