"""Loading helpers for the golden fixtures (tests/golden/*.npz)."""
import glob
import os

import numpy as np

from cases import CASES
from umans_amd.registry import pack_policy

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith(("example_", "csv_", "dynamic_", "api_")))


def example_names():
    return sorted(os.path.basename(p)[len("example_"):-4] for p in glob.glob(os.path.join(GOLDEN, "example_*.npz")))


def load_example(name):
    """(reference record, Scenario loaded from the XML by this library's loader)."""
    from umans_amd import gpu
    with np.load(os.path.join(GOLDEN, "example_" + name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    sc, meta = gpu.load_scenario(os.path.join(GOLDEN, "examples", name + ".xml"))
    return d, sc


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    agents = {k[len("agent_"):]: d[k] for k in d if k.startswith("agent_")}
    w = d["world"]
    if name.startswith("circle50"):
        packed = [pack_policy("global", [("ORCA", 1.0, {"timeHorizon": 5.0, "range": 8.0})])]
    else:
        case = name.rsplit("_w", 1)[0]
        if case.startswith("obst_"):
            case = case[len("obst_"):]
        m, costs, samp, relax = CASES[case]
        packed = [pack_policy(m, costs, sampling=samp, relaxation_time=relax)]
    return d, agents, packed, int(w[0]), float(w[1]), float(w[2]), float(w[3])


def neighbor_lists(d, step):
    """Reference neighbour lists of one step: list of (ids, phantom pos, float dsq)."""
    cnt = d["nb_count"]
    start = int(cnt[:step].sum())
    out = []
    o = start
    for c in cnt[step]:
        out.append((d["nb_id"][o:o + c], d["nb_pos"][o:o + c], d["nb_dsq"][o:o + c]))
        o += c
    return out


def state_at(agents, d, step):
    """Agent arrays with the reference state before `step` (0-based)."""
    a = dict(agents)
    for k in ("pos", "vel", "viewdir"):
        a[k] = d["state_" + k][step]
    return a
