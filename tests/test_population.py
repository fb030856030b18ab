"""Dynamic population and the reference's embedding API on the GPU path.

Fixtures (tests/golden/make_golden.py, from the compiled reference):
  * dynamic_Mixed.npz: the reference over examples/Dynamic-Mixed.xml (SocialForces and
    ORCA, scheduled insertions with equal start times, remove_at_goal, explicit and
    clashing IDs): the agent list (IDs in list order) and state after every step, and
    the CSV files its writer produced (WorldBase::DoStep, worldBase.cpp:79-128;
    CSVWriter, csvwriter.cpp:37-87); and the list of examples/Dynamic-Ids.xml after 25
    steps, whose scheduled agents share IDs.
  * api_Mixed.npz: the reference's own embedding API (src/Library/APIFunctions.cpp,
    built into oracle/_ref/libumans_ref_api.so) driven through
    embedding_api.api_sequence: what each call returned and the agents read back.
The GPU tests run the same scenario and the same calls through libumans_gpu.so."""
import os
import re

import numpy as np
import pytest

import embedding_api as E
from goldens import GOLDEN
from umans_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCENARIO = os.path.join(GOLDEN, "examples", "Dynamic-Mixed.xml")
API_NAMES = ["StartSimulation", "GetSimulationTimeStep", "DoSimulationSteps", "GetAgentPositions",
             "SetAgentPositions", "AddAgent", "RemoveAgent", "SetAgentGoal", "CleanUp"]


def load_states():
    d = np.load(os.path.join(GOLDEN, "dynamic_Mixed.npz"))
    off = np.concatenate([[0], np.cumsum(d["state_count"])])
    states = [{k: d["state_" + k][off[s]:off[s + 1]] for k in ("id", "pos", "vel", "viewdir")}
              for s in range(len(d["state_count"]))]
    return states, d


def test_embedding_api_header_and_exports():
    """include/umans_api.h declares the reference's nine functions (APIFunctions.h:54-112)
    and libumans_gpu.so exports them."""
    text = open(os.path.join(ROOT, "include", "umans_api.h")).read()
    for name in API_NAMES:
        assert re.search(r"\bbool\s+" + name + r"\s*\(", text), name
    L = E.bind_embedding_api(gpu.LIB_PATH)
    for name in API_NAMES:
        assert hasattr(L, name)


def test_embedding_api_without_simulation():
    """Every call returns false while no simulation runs (APIFunctions.cpp:68-205); a
    missing scenario file fails to start. No GPU is touched on these paths."""
    L = E.bind_embedding_api(gpu.LIB_PATH)
    import ctypes as C
    dt = C.c_float(0)
    assert not L.CleanUp()
    assert not L.DoSimulationSteps(1)
    assert not L.GetSimulationTimeStep(C.byref(dt))
    assert not L.RemoveAgent(0)
    assert not L.SetAgentGoal(0, 1.0, 1.0)
    assert E.get_agents(L) is None
    assert not L.StartSimulation(b"/nonexistent/scenario.xml", 1)


def test_scenario_ids_and_schedule_match_reference():
    """The loader gives the reference's IDs (clashing and explicit ones included) and
    start times; the agents present at time 0 are the reference's initial list."""
    states, _ = load_states()
    sc, meta = gpu.load_scenario(SCENARIO)
    now = meta["start_time"] <= 0
    np.testing.assert_array_equal(sc.agents["id"][now], states[0]["id"])
    np.testing.assert_array_equal(sc.agents["pos"][now], states[0]["pos"])
    every = set(np.concatenate([s["id"] for s in states]).tolist())
    assert set(sc.agents["id"].tolist()) == every


IDS_SCENARIO = os.path.join(GOLDEN, "examples", "Dynamic-Ids.xml")


def test_scheduled_agents_may_share_an_id():
    """A desired ID is refused only when an agent in the list holds it (GetAgent,
    worldBase.cpp:165): scheduled agents can end up sharing one. The loader gives the
    reference's IDs, and its start times order the insertions as the reference's list."""
    _, d = load_states()
    sc, meta = gpu.load_scenario(IDS_SCENARIO)
    assert sc.agents["id"].tolist() == [0, 1, 1, 1, 2]
    # the loader says that this backend keeps agents with one ID distinct
    assert any("Agent ID 1 is shared by 3 agents" in w for w in meta["warnings"])
    st = np.asarray(meta["start_time"])
    now = [i for i in range(len(st)) if st[i] <= 0]
    later = sorted((i for i in range(len(st)) if st[i] > 0), key=lambda i: st[i])
    assert [int(sc.agents["id"][i]) for i in now + later] == d["ids_list"].tolist()


@pytest.mark.gpu
def test_gpu_scheduled_insertion_order():
    _, d = load_states()
    w = gpu.World.from_scenario_file(IDS_SCENARIO)
    w.step(25)
    np.testing.assert_array_equal(w.records()[0], d["ids_list"])


@pytest.mark.gpu
def test_gpu_dynamic_population_follows_reference():
    """Each step starts from the reference's state (SetAgentPositions by ID); after the
    step the list — IDs in list order, after scheduled insertions and remove_at_goal
    removals — equals the reference's and the states agree to 1e-4."""
    states, _ = load_states()
    w = gpu.World.from_scenario_file(SCENARIO)
    assert w.num_scheduled() == 10
    for s in range(len(states) - 1):
        st = states[s]
        w.set_agent_states(st["id"], st["pos"], st["vel"], st["viewdir"])
        w.step(1)
        ids, pos, vel, view = w.records()
        ref = states[s + 1]
        np.testing.assert_array_equal(ids, ref["id"], err_msg=f"step {s + 1}")
        np.testing.assert_allclose(pos, ref["pos"], atol=1e-4, err_msg=f"step {s + 1}")
        np.testing.assert_allclose(vel, ref["vel"], atol=1e-4, err_msg=f"step {s + 1}")
    assert w.num_scheduled() == 0
    out = w.download()
    np.testing.assert_array_equal(out["id"], states[-1]["id"])


@pytest.mark.gpu
def test_gpu_dynamic_population_csv(tmp_path):
    """The CSV files of the free-running scenario: the same files, each with the
    reference's rows (agents log only while in the list), times exact, positions to
    1e-2 m after 60 free steps."""
    _, d = load_states()
    w = gpu.World.from_scenario_file(SCENARIO)
    w.csv_start(str(tmp_path))
    w.step(len(d["state_count"]) - 1)
    w.csv_stop()
    files = sorted(os.listdir(tmp_path), key=lambda f: int(f[len("output_"):-4]))
    assert [int(f[len("output_"):-4]) for f in files] == d["csv_id"].tolist()
    for f, ref in zip(files, d["csv_text"]):
        g = open(os.path.join(tmp_path, f)).read().splitlines()
        r = str(ref).splitlines()
        assert [ln.split(",")[0] for ln in g] == [ln.split(",")[0] for ln in r], f
        gv = np.array([[float(v) for v in ln.split(",")[1:]] for ln in g])
        rv = np.array([[float(v) for v in ln.split(",")[1:]] for ln in r])
        np.testing.assert_allclose(gv, rv, atol=1e-2, err_msg=f)


@pytest.mark.gpu
def test_gpu_population_edits_before_a_step():
    """Setters, additions and removals between steps are visible at once in the list
    (Agent::setGoal turns the agent to its goal; a removal moves the last row)."""
    w = gpu.World.from_scenario_file(SCENARIO)
    ids0 = w.records()[0]
    n0 = len(ids0)
    new = w.add_agents({"pos": np.array([[0.5, 3.0], [1.5, 2.0]], np.float32),
                        "id": np.array([3, 1000], np.int32),
                        "radius": np.full(2, 0.3, np.float32), "pref_speed": np.full(2, 1.3, np.float32),
                        "max_speed": np.full(2, 1.6, np.float32), "max_accel": np.full(2, 5.0, np.float32),
                        "policy": np.array([1, 0], np.int32)})
    assert new[1] == 1000 and new[0] not in ids0
    assert w.set_agent_goal(int(ids0[4]), 10.0, -3.0)
    assert w.remove_agent(int(ids0[2]))
    assert not w.remove_agent(123456)
    assert not w.set_agent_goal(123456, 0.0, 0.0)
    out = w.download()
    expect = list(ids0) + list(new)
    expect[2] = expect.pop()                  # the last row takes the removed one's place
    np.testing.assert_array_equal(out["id"], expect)
    r4 = expect.index(int(ids0[4]))
    np.testing.assert_array_equal(out["goal"][r4], [10.0, -3.0])
    dx, dy = np.float32(10.0) - out["pos"][r4][0], np.float32(-3.0) - out["pos"][r4][1]
    mag = np.sqrt(np.float32(dx * dx + dy * dy), dtype=np.float32)
    np.testing.assert_array_equal(out["viewdir"][r4], [dx / mag, dy / mag])
    np.testing.assert_array_equal(out["goal"][expect.index(1000)], [1.5, 2.0])
    assert w.n == n0 + 1
    w.step(3)
    assert w.n >= n0 - 10


@pytest.mark.gpu
def test_gpu_embedding_api_matches_reference():
    """The reference's embedding API over the GPU: the same return values, agent IDs and
    list order at every read-back, the edits exact before the next step, positions after
    5 steps to 1e-4 m and after 40 free steps to 1e-2 m."""
    ref = np.load(os.path.join(GOLDEN, "api_Mixed.npz"))
    L = E.bind_embedding_api(gpu.LIB_PATH)
    got = E.flatten(E.api_sequence(L, SCENARIO, 7))
    for k in ("started", "dt_ok", "steps_ok", "added", "goal_ok", "remove_ok", "set_ok", "cleanup",
              "after_cleanup", "a_id", "b0_id", "b_id", "c_id"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert got["dt"] == ref["dt"]
    np.testing.assert_allclose(got["a_pos"], ref["a_pos"], atol=1e-4)
    for k in ("b0_pos", "b0_vel", "b0_viewdir"):
        np.testing.assert_allclose(got[k], ref[k], atol=1e-4, err_msg=k)
    for k in ("b_pos", "c_pos"):
        np.testing.assert_allclose(got[k], ref[k], atol=1e-2, err_msg=k)


@pytest.mark.gpu
def test_gpu_embedding_api_unknown_policy():
    """AddAgent with a policy ID the scenario lacks returns false (the reference would
    add an agent without a policy)."""
    import ctypes as C
    L = E.bind_embedding_api(gpu.LIB_PATH)
    assert L.StartSimulation(SCENARIO.encode(), 1)
    rid = C.c_int(-1)
    assert not L.AddAgent(0.0, 0.0, 0.3, 1.3, 1.6, 5.0, 99, C.byref(rid), -1)
    assert L.AddAgent(0.0, 0.0, 0.3, 1.3, 1.6, 5.0, 0, C.byref(rid), -1)
    assert L.DoSimulationSteps(2)
    ids, pos, _, _ = E.get_agents(L)
    assert rid.value in ids.tolist()
    assert L.CleanUp()
