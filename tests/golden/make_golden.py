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


if __name__ == "__main__":
    main()
