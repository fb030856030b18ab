"""Write the golden fixtures under tests/golden/ from the compiled UMANS reference.

Run in a container that holds the reference sources (`make -C oracle ref` first):
    python tests/golden/make_golden.py
For every policy case (tests/golden/cases.py) and both world types it builds a small
dense random crowd (umans_amd.scenarios.random_cloud), loads the policy through the
reference's own XML loader, adds the agents through WorldBase::AddAgent, sets their
initial velocities, and records for each of STEPS calls of WorldBase::DoStep the
state before and after and every agent's neighbour list (ID, phantom position,
float squared distance). It also records the 50-agent circle crossing under ORCA.
The fixtures are data only (inputs and the reference's outputs).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import ref_py as R  # noqa: E402
from cases import CASES  # noqa: E402
from umans_amd import scenarios as S  # noqa: E402
from umans_amd.registry import to_reference_xml_policy  # noqa: E402

STEPS = 4
N = 64
SIZE = 12.0


def record(sim, steps):
    states, nbs = [sim.state()], []
    for _ in range(steps):
        sim.step(1)
        states.append(sim.state())
        nb = sim.neighbors()
        nbs.append(nb)
    out = {}
    for k in ("pos", "vel", "viewdir", "pref_vel"):
        out["state_" + k] = np.stack([s[k] for s in states]).astype(np.float32)
    cnt = np.array([[len(x[0]) for x in nb] for nb in nbs], np.int32)
    out["nb_count"] = cnt
    out["nb_id"] = np.concatenate([np.concatenate([x[0] for x in nb]) for nb in nbs]).astype(np.int32)
    out["nb_pos"] = np.concatenate([np.concatenate([x[1] for x in nb]) for nb in nbs]).astype(np.float32)
    out["nb_dsq"] = np.concatenate([np.concatenate([x[2] for x in nb]) for nb in nbs]).astype(np.float32)
    return out


def agent_arrays(sc):
    a = sc.agents
    return {"agent_" + k: np.asarray(v) for k, v in a.items()}


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libumans_ref.so missing: run `make -C oracle ref`")
    written = 0
    for world in (0, 1):
        for name, (method, costs, samp, relax) in CASES.items():
            xmlp = to_reference_xml_policy(method, costs, sampling=samp, relaxation_time=relax)
            sc = S.random_cloud(N, SIZE, [None], world_type=world, seed=11 + world)
            sim = R.RefSim(world, SIZE, SIZE, sc.dt, [xmlp], sc.agents, threads=1)
            data = record(sim, STEPS)
            sim.close()
            data.update(agent_arrays(sc))
            data["world"] = np.array([world, SIZE, SIZE, sc.dt], np.float64)
            np.savez_compressed(os.path.join(HERE, f"{name}_w{world}.npz"), **data)
            written += 1
    # circle crossing, 50 agents, ORCA (time horizon 5 s, range 8 m), 300 steps: keep
    # the full state every 10 steps and the neighbour lists of the first 3 steps.
    sc = S.circle(50)
    xmlp = to_reference_xml_policy("global", [("ORCA", 1.0, {"timeHorizon": 5.0, "range": 8.0})])
    sim = R.RefSim(0, 0, 0, sc.dt, [xmlp], sc.agents, threads=1)
    data = record(sim, 3)
    traj = [sim.state()]
    for _ in range(297):
        sim.step(1)
        traj.append(sim.state())
    sim.close()
    for k in ("pos", "vel"):
        data["traj_" + k] = np.stack([t[k] for t in traj[::9]]).astype(np.float32)
    data["traj_steps"] = np.arange(3, 301, 9, dtype=np.int32)
    data.update(agent_arrays(sc))
    data["world"] = np.array([0, 0, 0, sc.dt], np.float64)
    np.savez_compressed(os.path.join(HERE, "circle50_orca.npz"), **data)
    written += 1
    written += examples()
    written += csv_output()
    written += dynamic_population()
    written += obstacle_cases()
    print(f"wrote {written} fixtures to {HERE}")


EXAMPLE_FILES = ["Circle25-Dutra.xml", "Circle25-Karamouzas.xml", "Circle25-Moussaid.xml", "Circle25-ORCA.xml",
                 "Circle25-Paris.xml", "Circle25-PowerLaw.xml", "Circle25-RVO.xml", "Circle25-SocialForces.xml",
                 "Circle25-VanToll.xml", "Toric.xml"]
EXAMPLE_STEPS = 10


def examples():
    """Copy a few of the reference's example scenarios (XML inputs, with the policy and
    agent files they name) under tests/golden/examples/ and record the reference's state
    and neighbour lists over their first steps, loaded by its own FromConfigFile."""
    import re
    import shutil
    src = "/root/reference/examples"
    dst = os.path.join(HERE, "examples")
    os.makedirs(os.path.join(dst, "policies"), exist_ok=True)
    os.makedirs(os.path.join(dst, "agents"), exist_ok=True)
    n = 0
    for f in EXAMPLE_FILES:
        shutil.copy(os.path.join(src, f), os.path.join(dst, f))
        for ref in re.findall(r'file="([^"]+)"', open(os.path.join(src, f)).read()):
            shutil.copy(os.path.join(src, ref), os.path.join(dst, ref))
        sim = R.RefSim.from_file(os.path.join(dst, f))
        data = record(sim, EXAMPLE_STEPS)
        sim.close()
        np.savez_compressed(os.path.join(HERE, "example_" + f[:-4] + ".npz"), **data)
        n += 1
    return n


CSV_EXAMPLE = "Circle25-ORCA.xml"
CSV_STEPS = 10


def csv_output():
    """Run the reference's CSV writer (CrowdSimulator::StartCSVOutput, flushed when the
    simulator is destroyed) over the first steps of an example and keep every agent's
    output_<id>.csv text, with the states after each step."""
    import tempfile
    path = os.path.join(HERE, "examples", CSV_EXAMPLE)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "csv")
        os.makedirs(out)      # CreateDirectoryIfNonExistent cannot create absolute paths
        sim = R.RefSim.from_file(path)
        sim.start_csv(out, False)
        states = [sim.state()]
        for _ in range(CSV_STEPS):
            sim.step(1)
            states.append(sim.state())
        sim.close()
        ids = states[0]["id"]
        text = np.array([open(os.path.join(out, f"output_{i}.csv")).read() for i in ids])
    np.savez_compressed(os.path.join(HERE, "csv_" + CSV_EXAMPLE[:-4] + ".npz"), id=ids, text=text,
                        state_pos=np.stack([s["pos"] for s in states]).astype(np.float32),
                        state_id=np.stack([s["id"] for s in states]).astype(np.int32))
    return 1


DYNAMIC_XML = "Dynamic-Mixed.xml"
DYNAMIC_STEPS = 60


def dynamic_scenario_xml():
    """Two groups crossing under two policies (XML ids 0 and 7), some agents removed at
    their goal, some joining later (equal start times included), some with explicit IDs
    or an ID already in the list. No two agents share an ID: the reference resolves
    neighbours by ID (WorldBase::GetAgent_const, worldBase.cpp:51-60), so agents sharing
    one see the same neighbour twice there."""
    rng = np.random.default_rng(5)
    pol = ("  <Policies>\n"
           "    <Policy id=\"0\" OptimizationMethod=\"gradient\" RelaxationTime=\"0.5\">\n"
           "      <costfunction range=\"5\" name=\"SocialForcesAvoidance\" />\n"
           "      <costfunction name=\"GoalReachingForce\" />\n"
           "    </Policy>\n"
           "    <Policy id=\"7\" OptimizationMethod=\"global\" RelaxationTime=\"0\">\n"
           "      <costfunction range=\"5\" name=\"ORCA\" />\n"
           "    </Policy>\n"
           "  </Policies>\n")
    agents = []

    def agent(x, y, gx, gy, policy, attrs=""):
        agents.append(f'    <Agent rad="0.3" pref_speed="1.3" max_speed="1.6"{attrs}>\n'
                      f'      <pos x="{x:.3f}" y="{y:.3f}"/>\n      <goal x="{gx:.3f}" y="{gy:.3f}"/>\n'
                      f'      <Policy id="{policy}"/>\n    </Agent>\n')
    for i in range(16):
        x, y = -6.0 + (i % 4) * 0.8 + rng.uniform(-0.1, 0.1), -1.5 + (i // 4) * 0.9 + rng.uniform(-0.1, 0.1)
        gx = 5.0 if i % 2 else -2.5
        agent(x, y, gx, y + rng.uniform(-0.5, 0.5), 0, ' remove_at_goal="true"' if i % 2 == 0 else "")
    for i in range(16):
        x, y = 6.0 - (i % 4) * 0.8 + rng.uniform(-0.1, 0.1), -1.4 + (i // 4) * 0.9 + rng.uniform(-0.1, 0.1)
        gx = -5.0 if i % 3 else 2.5
        agent(x, y, gx, y + rng.uniform(-0.5, 0.5), 7, ' remove_at_goal="true"' if i % 3 == 0 else "")
    late = [(0.5, ""), (0.5, ' id="3"'), (1.0, ' id="500"'), (1.05, ""), (2.0, ' remove_at_goal="true"'),
            (2.0, ""), (2.0, ' id="40"'), (3.3, ""), (0.5, ' id="700"'), (4.0, ' remove_at_goal="true"')]
    for j, (t, extra) in enumerate(late):
        side = -1 if j % 2 == 0 else 1
        x, y = side * 7.0, -2.0 + 0.45 * j
        agent(x, y, -side * 4.0, y, 0 if side < 0 else 7, f' start_time="{t}"' + extra)
    return ('<?xml version="1.0" encoding="utf-8"?>\n<Simulation delta_time="0.1" end_time="10">\n'
            '  <World type="Infinite" />\n' + pol + "  <Agents>\n" + "".join(agents) + "  </Agents>\n</Simulation>\n")


def ragged(states, keys=("id", "pos", "vel", "viewdir")):
    out = {"count": np.array([len(s["id"]) for s in states], np.int32)}
    for k in keys:
        out[k] = np.concatenate([s[k] for s in states])
    return out


def dynamic_population():
    """The reference over a scenario with scheduled insertions and remove_at_goal: the
    agent list (IDs in list order) and states after every step; and the reference's
    embedding API (src/Library/APIFunctions.cpp) driven through a fixed call sequence."""
    sys.path.insert(0, os.path.dirname(HERE))
    import embedding_api as E
    path = os.path.join(HERE, "examples", DYNAMIC_XML)
    with open(path, "w") as f:
        f.write(dynamic_scenario_xml())
    # IDs only: a scheduled agent may be given the ID of another scheduled agent
    # (GetAgent looks at the list only, worldBase.cpp:165), so two agents share an ID
    ids_path = os.path.join(HERE, "examples", "Dynamic-Ids.xml")
    with open(ids_path, "w") as f:
        f.write('<?xml version="1.0" encoding="utf-8"?>\n<Simulation delta_time="0.1">\n  <World type="Infinite" />\n'
                '  <Policies file="policies/SocialForces.xml" />\n  <Agents>\n'
                + "".join(f'    <Agent{a}>\n      <pos x="{x}" y="0"/>\n      <goal x="{x}" y="5"/>\n'
                          f'      <Policy id="0"/>\n    </Agent>\n'
                          for x, a in ((0, ""), (3, ' start_time="1"'), (6, ' start_time="2" id="1"'),
                                       (9, ' id="1"'), (12, ' start_time="0.5"')))
                + "  </Agents>\n</Simulation>\n")
    sim = R.RefSim.from_file(ids_path)
    sim.step(25)
    ids_list = sim.state()["id"]
    sim.close()
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        sim = R.RefSim.from_file(path)
        sim.start_csv(tmp, False)
        states = [sim.state()]
        for _ in range(DYNAMIC_STEPS):
            sim.step(1)
            states.append(sim.state())
        sim.close()
        files = sorted(os.listdir(tmp), key=lambda f: int(f[len("output_"):-4]))
        csv_ids = np.array([int(f[len("output_"):-4]) for f in files], np.int32)
        csv_text = np.array([open(os.path.join(tmp, f)).read() for f in files])
    np.savez_compressed(os.path.join(HERE, "dynamic_Mixed.npz"), csv_id=csv_ids, csv_text=csv_text, ids_list=ids_list,
                        **{"state_" + k: v for k, v in ragged(states).items()})
    L = E.bind_embedding_api(R.API_LIB_PATH)
    rec = E.api_sequence(L, path, 7)
    np.savez_compressed(os.path.join(HERE, "api_Mixed.npz"), **E.flatten(rec))
    return 2


def obstacle_cases():
    """The policy cases of OBSTACLE_CASES in a world with the OBSTACLES polygons, both
    world types: states and neighbour lists over STEPS reference steps, plus the edges."""
    from cases import OBSTACLE_CASES, OBSTACLES
    import oracle_py as O
    edges = O.polygon_edges([len(p) for p in OBSTACLES], [xy for p in OBSTACLES for xy in p])
    n = 0
    for world in (0, 1):
        for name in OBSTACLE_CASES:
            method, costs, samp, relax = CASES[name]
            xmlp = to_reference_xml_policy(method, costs, sampling=samp, relaxation_time=relax)
            sc = S.random_cloud(N, SIZE, [None], world_type=world, seed=23 + world)
            sim = R.RefSim(world, SIZE, SIZE, sc.dt, [xmlp], sc.agents, threads=1, obstacles=OBSTACLES)
            data = record(sim, STEPS)
            sim.close()
            data.update(agent_arrays(sc))
            data["world"] = np.array([world, SIZE, SIZE, sc.dt], np.float64)
            data["obstacle_edges"] = edges
            np.savez_compressed(os.path.join(HERE, f"obst_{name}_w{world}.npz"), **data)
            n += 1
    return n


if __name__ == "__main__":
    main()
