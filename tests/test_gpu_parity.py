"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

* golden fixtures (written by the compiled reference): from the reference's state,
  one GPU step reproduces the reference's neighbour lists exactly (IDs and order) and
  its next velocities, positions and viewing directions;
* the FP64 oracle: new velocities within 1e-5 relative error (|dv| / max(|v|, 1e-3)),
  except where the reference's own FP32 result departs from FP64 in the same way
  (a discrete choice that FP32 rounding decides), in which case the GPU must equal the
  FP32 result;
* neighbour lists bit-exact against the oracle at 10k agents with and without a cap,
  in infinite and toric worlds (query disks crossing the seams);
* the obst_* fixtures repeat the cases whose models read obstacle edges in a world with
  polygons (tests/golden/cases.py OBSTACLES), toric seams included.
"""
import numpy as np
import pytest

import oracle_py as O
from goldens import example_names, fixture_names, load, load_example, neighbor_lists, state_at
from cases import OBSTACLES
from gpu_util import REL_TOL, gpu_step, rel_err
from umans_amd import scenarios as S

pytestmark = pytest.mark.gpu

NAMES = [n for n in fixture_names()]


@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_reference_fixture(name):
    d, agents, packed, world, w, h, dt = load(name)
    steps = d["nb_count"].shape[0]
    rng = np.array([O.rng_seed(i) for i in agents["id"]], np.uint32)
    obstacles = OBSTACLES if name.startswith("obst_") else None
    for s in range(steps):
        a = state_at(agents, d, s)
        a["rng"] = rng
        out, (cnt, ids, d2, seg), st = gpu_step(world, w, h, dt, packed, a, obstacles=obstacles)
        ref_nb = neighbor_lists(d, s)
        for i, (rid, _, _) in enumerate(ref_nb):
            assert cnt[i] == len(rid), (name, s, i)
            np.testing.assert_array_equal(ids[i, :len(rid)], rid, err_msg=f"{name} step {s} agent {i}")
        # velocities against the reference FP32 step and the FP64 oracle
        o64 = O.step(O.make_world(world, w, h, dt, 0, edges=d.get("obstacle_edges")), packed, a, precision=1)
        ref32 = d["state_vel"][s + 1]
        e64 = rel_err(out["vel"], o64["vel64"])
        e32 = rel_err(out["vel"], ref32)
        ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
        assert ok.all(), f"{name} step {s}: agents {np.where(~ok)[0][:8]} e64 {e64[~ok][:4]} e32 {e32[~ok][:4]}"
        np.testing.assert_allclose(out["pos"], d["state_pos"][s + 1], rtol=0, atol=1e-5)
        rng = out["rng"]


def test_gpu_circle_trajectory():
    """50-agent circle crossing under ORCA: the free-running GPU trajectory follows the
    reference's over 300 steps."""
    d, agents, packed, world, w, h, dt = load("circle50_orca")
    a = state_at(agents, d, 0)
    out, _, _ = gpu_step(world, w, h, dt, packed, a, steps=3, record=False)
    np.testing.assert_allclose(out["pos"], d["traj_pos"][0], atol=1e-5)
    W = __import__("umans_amd.gpu", fromlist=["World"]).World(world, w, h, dt, packed)
    W.upload(a)
    k = 0
    for s in range(1, 301):
        W.step(1)
        if s in set(d["traj_steps"].tolist()):
            o = W.download()
            np.testing.assert_allclose(o["pos"], d["traj_pos"][k], atol=1e-3, err_msg=f"step {s}")
            k += 1
    W.close()


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("cap", [16, 0])
def test_neighbor_lists_bit_exact_10k(world_type, cap):
    pol = S.powerlaw_sampling_policy(rng=5.0)
    sc = S.random_cloud(10_000, 100.0, [pol], world_type=world_type, seed=21 + cap)
    a = sc.agents
    out, (cnt, ids, d2, seg), st = gpu_step(world_type, sc.width, sc.height, sc.dt, [pol], a, max_neighbors=cap)
    stride = 128
    o = O.step(O.make_world(world_type, sc.width, sc.height, sc.dt, cap), [pol], a, precision=0, list_stride=stride)
    np.testing.assert_array_equal(cnt, o["nb_count"])
    for i in range(sc.n):
        c = cnt[i]
        np.testing.assert_array_equal(ids[i, :c], o["nb_id"][i, :c], err_msg=f"agent {i}")
        np.testing.assert_array_equal(d2[i, :c], o["nb_d2"][i, :c])
        np.testing.assert_array_equal(seg[i, :c], o["nb_seg"][i, :c])
    if world_type == 1 and cap == 0:
        assert (o["nb_seg"] > 0).any()   # query disks crossed the seams
    if cap:
        assert st["overflow_agents"] == int((o["nb_full"] > cap).sum())


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("cap", [12, 16])
def test_orca_capped_lists_dense(world_type, cap):
    """ORCA with lists of at most 12 or 16 neighbours keeps its half-planes in registers
    (the longer lists of the other tests use the solver over memory). A crowd dense
    enough that agents overlap and many linear programs are infeasible (linearProgram3)
    against the FP32 and FP64 oracles."""
    pol = S.orca_policy(time_horizon=5.0, rng=6.0)
    sc = S.random_cloud(4000, 40.0, [pol], world_type=world_type, max_neighbors=cap, seed=70 + cap, speed=1.5)
    a = sc.agents
    out, _, st = gpu_step(world_type, sc.width, sc.height, sc.dt, [pol], a, cap, record=False)
    wd = O.make_world(world_type, sc.width, sc.height, sc.dt, cap)
    o32 = O.step(wd, [pol], a, precision=0)
    o64 = O.step(wd, [pol], a, precision=1)
    e64 = rel_err(out["vel"], o64["vel64"])
    e32 = rel_err(out["vel"], o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    assert ok.all(), f"{(~ok).sum()} agents off"
    assert st["overflow_agents"] > 0      # the lists are full: the cap is what is tested


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("case,cap", [("karamouzas_sampling", 16), ("moussaid_sampling", 16),
                                      ("paris_sampling", 12), ("vantoll_sampling", 16),
                                      ("ttcadca_sampling", 12)])
def test_cooperative_sampling_capped_lists(world_type, case, cap):
    """Regular sampling over per-candidate neighbour costs with lists of at most 12 or 16
    entries runs four waves per 64 agents with the neighbours staged in LDS (longer lists
    use one wave and read them from memory): one step of a dense cloud against the FP32
    and FP64 oracles."""
    from cases import CASES
    from umans_amd.registry import pack_policy
    m, costs, samp, relax = CASES[case]
    pol = pack_policy(m, costs, sampling=samp, relaxation_time=relax)
    sc = S.random_cloud(3000, 45.0, [pol], world_type=world_type, max_neighbors=cap, seed=90 + cap)
    a = sc.agents
    out, _, st = gpu_step(world_type, sc.width, sc.height, sc.dt, [pol], a, cap, record=False)
    wd = O.make_world(world_type, sc.width, sc.height, sc.dt, cap)
    o32 = O.step(wd, [pol], a, precision=0)
    o64 = O.step(wd, [pol], a, precision=1)
    e64 = rel_err(out["vel"], o64["vel64"])
    e32 = rel_err(out["vel"], o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    assert ok.all(), f"{(~ok).sum()} agents off"
    assert st["overflow_agents"] > 0


def test_mixed_crowd_100k_one_step():
    """The mixed crowd (PowerLaw sampling / SocialForces gradient / ORCA) at 100k agents
    and the same density: one step against the FP32 and FP64 oracles."""
    sc = S.mixed(n=100_000, size=632.0)
    a = sc.agents
    out, _, st = gpu_step(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, a, sc.max_neighbors, record=False)
    wd = O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors)
    o32 = O.step(wd, sc.policies, a, precision=0, nthreads=8)
    o64 = O.step(wd, sc.policies, a, precision=1, nthreads=8)
    e64 = rel_err(out["vel"], o64["vel64"])
    e32 = rel_err(out["vel"], o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    assert ok.mean() == 1.0, f"{(~ok).sum()} agents off: e64 {np.sort(e64)[-5:]}"
    np.testing.assert_array_equal(out["id"], a["id"])


@pytest.mark.parametrize("dense", [False, True])
def test_force_sampling_ranking_and_ties(dense):
    """Regular sampling of the force-model costs (PowerLaw + GoalReachingForce,
    costFunction.cpp:66-164) on crowds built to tie: agents standing at their goals with
    zero or tiny velocities (targets at the centre of a ring of equal candidates, decided
    by the first candidate in loop order), agents walking straight at their goals (a grid
    symmetric about the target), and random ones; isolated (no neighbours, so no
    transcendental in the costs) or dense. The kernel ranks candidates by
    sum(coeff |v - target|^2) and leaves near-ties to the exact cost loop: every chosen
    velocity must equal the FP32 oracle's (a different candidate is off by a sampling
    step, 0.3 m/s)."""
    pol = S.powerlaw_sampling_policy()
    n = 4000
    r = np.random.default_rng(77 + dense)
    a = S._agents(n)
    side = int(np.ceil(np.sqrt(n)))
    spacing = 1.2 if dense else 20.0
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)[:n]
    a["pos"][:] = ((g - side / 2) * spacing + r.uniform(-0.1, 0.1, (n, 2))).astype(np.float32)
    grp = np.arange(n) % 4
    a["goal"][:] = a["pos"]                                             # groups 0, 1: at the goal
    a["vel"][grp == 1] = r.uniform(-1e-3, 1e-3, ((grp == 1).sum(), 2)).astype(np.float32)
    d = r.normal(size=(n, 2)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    walk = grp == 2                                                     # straight at the goal
    a["goal"][walk] = (a["pos"][walk] + 30.0 * d[walk]).astype(np.float32)
    a["vel"][walk] = (1.2 * d[walk]).astype(np.float32)
    rnd = grp == 3
    a["goal"][rnd] = (a["pos"][rnd] + r.uniform(-20, 20, (rnd.sum(), 2))).astype(np.float32)
    a["vel"][rnd] = r.uniform(-1.5, 1.5, (rnd.sum(), 2)).astype(np.float32)
    a["viewdir"] = S.initial_viewdir(a["pos"], a["goal"])
    out, _, _ = gpu_step(0, 0.0, 0.0, 0.1, [pol], a, max_neighbors=12, record=False)
    wd = O.make_world(0, 0.0, 0.0, 0.1, 12)
    o32 = O.step(wd, [pol], a, precision=0, nthreads=8)
    o64 = O.step(wd, [pol], a, precision=1, nthreads=8)
    e32 = rel_err(out["vel"], o32["vel"])
    e64 = rel_err(out["vel"], o64["vel64"])
    if dense:
        # PowerLaw's pair force calls expf: the chosen candidate may differ from the FP32
        # reference's only where the FP32 and FP64 choices differ too
        ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    else:
        ok = e32 <= 1e-6
    assert ok.all(), f"agents {np.where(~ok)[0][:8]} (groups {grp[~ok][:8]}): e32 {e32[~ok][:4]} e64 {e64[~ok][:4]}"


@pytest.mark.parametrize("name", example_names())
def test_gpu_on_reference_examples(name):
    """Example scenarios of the reference read by the XML loader and stepped on the GPU
    from the reference's state each step: neighbour sets exact (order may differ only
    among equal distances), velocities to the parity bar."""
    from test_oracle_golden import tie_equal
    d, sc = load_example(name)
    a = dict(sc.agents)
    for s in range(d["nb_count"].shape[0]):
        for k in ("pos", "vel", "viewdir"):
            a[k] = d["state_" + k][s]
        out, (cnt, ids, d2, seg), st = gpu_step(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, a)
        a["rng"] = out["rng"]
        for i, (rid, _, _) in enumerate(neighbor_lists(d, s)):
            assert tie_equal(rid, ids[i, :cnt[i]], d2[i]), (name, s, i)
        o64 = O.step(O.make_world(sc.world_type, sc.width, sc.height, sc.dt, 0), sc.policies, a, precision=1)
        e64 = rel_err(out["vel"], o64["vel64"])
        e32 = rel_err(out["vel"], d["state_vel"][s + 1])
        ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
        assert ok.all(), (name, s, np.where(~ok)[0][:5])
