"""Pins the CPU oracle to the compiled reference: for every golden fixture, one oracle
DoStep from the reference's state reproduces the reference's next state bit for bit
(FP32 mode) and its neighbour lists exactly (IDs and order)."""
import numpy as np
import pytest

import oracle_py as O
from goldens import example_names, fixture_names, load, load_example, neighbor_lists, state_at

NAMES = fixture_names()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference(name):
    d, agents, packed, world, w, h, dt = load(name)
    wd = O.make_world(world, w, h, dt, 0, edges=d.get("obstacle_edges"))
    rng = np.array([O.rng_seed(i) for i in agents["id"]], np.uint32)
    steps = d["nb_count"].shape[0]
    for s in range(steps):
        a = state_at(agents, d, s)
        a["rng"] = rng
        out = O.step(wd, packed, a, precision=0, list_stride=128)
        rng = out["rng"]
        ref_nb = neighbor_lists(d, s)
        for i, (ids, pos, dsq) in enumerate(ref_nb):
            assert out["nb_count"][i] == len(ids)
            np.testing.assert_array_equal(out["nb_id"][i, :len(ids)], ids)
        np.testing.assert_array_equal(out["vel"], d["state_vel"][s + 1])
        np.testing.assert_array_equal(out["pos"], d["state_pos"][s + 1])
        np.testing.assert_array_equal(out["viewdir"], d["state_viewdir"][s + 1])
        np.testing.assert_array_equal(out["pref_vel"], d["state_pref_vel"][s + 1])


def test_oracle_circle_trajectory():
    """Free-running oracle on the 50-agent circle crossing stays on the reference
    trajectory (ORCA, 300 steps)."""
    d, agents, packed, world, w, h, dt = load("circle50_orca")
    wd = O.make_world(world, w, h, dt, 0)
    a = state_at(agents, d, 0)
    steps = set(d["traj_steps"].tolist())
    k = 0
    for s in range(1, 301):
        out = O.step(wd, packed, a, precision=0)
        a["pos"], a["vel"], a["viewdir"] = out["pos"], out["vel"], out["viewdir"]
        if s in steps:
            np.testing.assert_array_equal(a["pos"], d["traj_pos"][k])
            k += 1
    assert k == len(d["traj_steps"])


def tie_equal(ref_ids, ids, d2):
    """The same neighbours in an order that differs at most among equal distances."""
    if np.array_equal(ref_ids, ids):
        return True
    if sorted(ref_ids.tolist()) != sorted(ids.tolist()):
        return False
    at = {j: k for k, j in enumerate(ids)}
    return all(d2[at[j]] == d2[k] for k, j in enumerate(ref_ids))


@pytest.mark.parametrize("name", example_names())
def test_oracle_on_reference_examples(name):
    """The reference's example scenarios, loaded by this library's XML loader: the oracle
    reproduces the reference's steps (started from the reference's state each step)."""
    d, sc = load_example(name)
    wd = O.make_world(sc.world_type, sc.width, sc.height, sc.dt, 0)
    a = dict(sc.agents)
    for s in range(d["nb_count"].shape[0]):
        for k in ("pos", "vel", "viewdir"):
            a[k] = d["state_" + k][s]
        out = O.step(wd, sc.policies, a, precision=0, list_stride=128)
        a["rng"] = out["rng"]
        for i, (rid, _, _) in enumerate(neighbor_lists(d, s)):
            assert tie_equal(rid, out["nb_id"][i, :out["nb_count"][i]], out["nb_d2"][i])
        np.testing.assert_allclose(out["vel"], d["state_vel"][s + 1], rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(out["pos"], d["state_pos"][s + 1], rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("obst_")])
def test_obstacle_fixtures_depend_on_obstacles(name):
    """The obstacle fixtures exercise the obstacle code: the same step without the edges
    gives a different result for some agents (so the bit-exact match above pins the
    obstacle search, the segment time-to-collision, SocialForces' obstacle force and the
    obstacle contact forces, whichever the case uses)."""
    d, agents, packed, world, w, h, dt = load(name)
    wd = O.make_world(world, w, h, dt, 0)
    a = state_at(agents, d, 0)
    a["rng"] = np.array([O.rng_seed(i) for i in agents["id"]], np.uint32)
    out = O.step(wd, packed, a, precision=0)
    assert (out["vel"] != d["state_vel"][1]).any(axis=1).sum() >= 3


@pytest.mark.parametrize("name", [n for n in NAMES if "random" in n])
def test_oracle_default_engine_state_is_the_reference_seed(name):
    """Given no engine state, the oracle seeds each agent's engine with its ID, as
    Agent::Agent does (agent.cpp:38): the reference's first step of a random-sampling
    fixture is reproduced bit for bit with no rng array, and not with a planted
    state of 1 for every agent."""
    d, agents, packed, world, w, h, dt = load(name)
    wd = O.make_world(world, w, h, dt, 0, edges=d.get("obstacle_edges"))
    a = state_at(agents, d, 0)
    a.pop("rng", None)
    out = O.step(wd, packed, a, precision=0)
    np.testing.assert_array_equal(out["vel"], d["state_vel"][1])
    a["rng"] = np.ones(len(a["id"]), np.uint32)
    planted = O.step(wd, packed, a, precision=0)
    assert not np.array_equal(planted["vel"], d["state_vel"][1])
