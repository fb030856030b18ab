"""ctypes binding of the reference's embedding API (APIFunctions.h:30-113), for any
library that exports it: the reference's own build (oracle/_ref/libumans_ref_api.so,
test infrastructure) or umans_amd/libumans_gpu.so. C++ references are passed as
pointers; AddAgent's default argument is passed explicitly."""
import ctypes as C

import numpy as np


class AgentData(C.Structure):
    """APIFunctions.h:33-49."""
    _fields_ = [("id", C.c_int), ("position_x", C.c_float), ("position_y", C.c_float),
                ("velocity_x", C.c_float), ("velocity_y", C.c_float),
                ("viewingDirection_x", C.c_float), ("viewingDirection_y", C.c_float)]


def bind_embedding_api(path):
    L = C.CDLL(path)
    L.StartSimulation.argtypes = [C.c_char_p, C.c_int]
    L.StartSimulation.restype = C.c_bool
    L.GetSimulationTimeStep.argtypes = [C.POINTER(C.c_float)]
    L.GetSimulationTimeStep.restype = C.c_bool
    L.DoSimulationSteps.argtypes = [C.c_int]
    L.DoSimulationSteps.restype = C.c_bool
    L.GetAgentPositions.argtypes = [C.POINTER(C.POINTER(AgentData)), C.POINTER(C.c_int)]
    L.GetAgentPositions.restype = C.c_bool
    L.SetAgentPositions.argtypes = [C.POINTER(AgentData), C.c_int]
    L.SetAgentPositions.restype = C.c_bool
    L.AddAgent.argtypes = [C.c_float] * 6 + [C.c_int, C.POINTER(C.c_int), C.c_int]
    L.AddAgent.restype = C.c_bool
    L.RemoveAgent.argtypes = [C.c_int]
    L.RemoveAgent.restype = C.c_bool
    L.SetAgentGoal.argtypes = [C.c_int, C.c_float, C.c_float]
    L.SetAgentGoal.restype = C.c_bool
    L.CleanUp.argtypes = []
    L.CleanUp.restype = C.c_bool
    return L


def get_agents(L):
    """GetAgentPositions as arrays (ids, pos, vel, viewdir), or None on failure."""
    ptr = C.POINTER(AgentData)()
    n = C.c_int(0)
    if not L.GetAgentPositions(C.byref(ptr), C.byref(n)):
        return None
    raw = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(max(n.value, 1) * 7,))[:n.value * 7]
    raw = raw.reshape(n.value, 7).copy()
    return raw[:, 0].view(np.int32).copy(), raw[:, 1:3].copy(), raw[:, 3:5].copy(), raw[:, 5:7].copy()


def snapshot(L):
    r = get_agents(L)
    return {"id": r[0], "pos": r[1], "vel": r[2], "viewdir": r[3]}


def api_sequence(L, scenario_path, policy_id):
    """The calls an embedding application makes, in order; returns what each call
    returned and the agents read back at three points. policy_id: an XML policy id of
    the scenario, used for added agents."""
    out = {"started": bool(L.StartSimulation(str(scenario_path).encode(), 1))}
    dt = C.c_float(0)
    out["dt_ok"] = bool(L.GetSimulationTimeStep(C.byref(dt)))
    out["dt"] = dt.value
    out["steps_ok"] = [bool(L.DoSimulationSteps(5))]
    out["a"] = snapshot(L)
    added = []
    for x, y, desired in ((0.5, 3.0, 3), (-0.5, -3.0, -1), (1.5, 2.0, 1000)):
        rid = C.c_int(-7)
        ok = L.AddAgent(x, y, 0.3, 1.3, 1.6, 5.0, policy_id, C.byref(rid), desired)
        added.append((bool(ok), rid.value))
    out["added"] = added
    ids = [a[1] for a in added]
    out["goal_ok"] = [bool(L.SetAgentGoal(ids[0], 4.0, 3.0)), bool(L.SetAgentGoal(ids[2], -4.0, 2.0)),
                      bool(L.SetAgentGoal(123456, 1.0, 1.0))]
    out["remove_ok"] = [bool(L.RemoveAgent(2)), bool(L.RemoveAgent(123456)), bool(L.RemoveAgent(ids[1]))]
    recs = (AgentData * 3)()
    for r, (aid, px, py, vx, vy, dx, dy) in zip(recs, ((5, -3.0, 0.25, 0.5, 0.0, 1.0, 0.0),
                                                        (ids[2], 1.25, 2.5, -0.25, 0.0, -1.0, 0.0),
                                                        (999999, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0))):
        r.id, r.position_x, r.position_y, r.velocity_x, r.velocity_y = aid, px, py, vx, vy
        r.viewingDirection_x, r.viewingDirection_y = dx, dy
    out["set_ok"] = bool(L.SetAgentPositions(recs, 3))
    out["b0"] = snapshot(L)                  # the changes, before any step
    out["steps_ok"].append(bool(L.DoSimulationSteps(10)))
    out["b"] = snapshot(L)
    out["steps_ok"].append(bool(L.DoSimulationSteps(25)))
    out["c"] = snapshot(L)
    out["cleanup"] = [bool(L.CleanUp()), bool(L.CleanUp())]
    out["after_cleanup"] = [bool(L.DoSimulationSteps(1)), bool(L.GetSimulationTimeStep(C.byref(dt)))]
    return out


def flatten(rec):
    """The record of api_sequence as flat arrays (for an .npz fixture)."""
    flat = {}
    for k, v in rec.items():
        if isinstance(v, dict):
            for kk, vv in v.items():
                flat[f"{k}_{kk}"] = np.asarray(vv)
        else:
            flat[k] = np.asarray(v)
    return flat
