"""ctypes binding of the compiled UMANS reference (oracle/_ref/libumans_ref.so) — TEST
INFRASTRUCTURE ONLY.

The library is compiled from the reference's own sources by `make -C oracle ref`
(only where those sources exist); oracle/ref_driver.cpp exposes CrowdSimulator
loading, WorldBase::AddAgent, DoStep and state/neighbour read-back. Tests use it
to pin the oracle and to write golden fixtures; bench.py may time it as the CPU
baseline. It never participates in the product path.
"""
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libumans_ref.so")
API_LIB_PATH = os.path.join(HERE, "_ref", "libumans_ref_api.so")

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        L.ref_load.argtypes = [C.c_char_p]; L.ref_load.restype = C.c_void_p
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.ref_num_agents.argtypes = [C.c_void_p]; L.ref_num_agents.restype = C.c_int
        L.ref_add_agents.argtypes = [C.c_void_p, C.c_int, _I, _F, _F, _F, _F, _F, _F, _I]
        L.ref_add_agents.restype = C.c_int
        L.ref_set_velocities.argtypes = [C.c_void_p, C.c_int, _F, _F]
        L.ref_step.argtypes = [C.c_void_p, C.c_int]
        L.ref_start_csv.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.ref_get_settings.argtypes = [C.c_void_p, _F, _F, _F, _F, _I, C.c_int]
        L.ref_time_steps.argtypes = [C.c_void_p, C.c_int]; L.ref_time_steps.restype = C.c_double
        L.ref_get_state.argtypes = [C.c_void_p, _I, _F, _F, _F, _F, _F]
        L.ref_total_neighbors.argtypes = [C.c_void_p]; L.ref_total_neighbors.restype = C.c_longlong
        L.ref_get_neighbors.argtypes = [C.c_void_p, _I, _I, _F, _F]
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


METHOD_NAMES = {0: "gradient", 1: "sampling", 2: "global"}
TYPE_NAMES = {0: "regular", 1: "random"}
BASE_NAMES = {0: "zero", 1: "current velocity"}
DIR_NAMES = {0: "unit", 1: "current velocity", 2: "preferred velocity"}
RADIUS_NAMES = {0: "preferred speed", 1: "maximum speed", 2: "maximum acceleration"}


def policy_xml(pid, pol):
    """pol: dict with 'method', optional 'sampling' attrs, 'relaxation_time',
    'contact_force_scale', and 'costs' = [(name, {attr: value})]."""
    attrs = [f'id="{pid}"', f'OptimizationMethod="{pol["method"]}"']
    s = pol.get("sampling")
    if s:
        for k, v in s.items():
            attrs.append(f'{k}="{v}"')
    if "relaxation_time" in pol:
        attrs.append(f'RelaxationTime="{pol["relaxation_time"]}"')
    if "contact_force_scale" in pol:
        attrs.append(f'ContactForceScale="{pol["contact_force_scale"]}"')
    body = []
    for name, params in pol["costs"]:
        pa = " ".join(f'{k}="{v!r}"' if isinstance(v, float) else f'{k}="{v}"' for k, v in params.items())
        body.append(f'    <costfunction name="{name}" {pa} />')
    return "  <Policy " + " ".join(attrs) + ">\n" + "\n".join(body) + "\n  </Policy>"


def obstacles_xml(polygons):
    """<Obstacles> block (crowdSimulator.cpp:496-514): one <Obstacle> of <Point>s each."""
    out = []
    for poly in polygons:
        pts = "".join('<Point x="%r" y="%r"/>' % (float(x), float(y)) for x, y in poly)
        out.append("<Obstacle>" + pts + "</Obstacle>")
    return "<Obstacles>" + "".join(out) + "</Obstacles>"


def scenario_xml(world_type, width, height, dt, policies, obstacles=None):
    inner = obstacles_xml(obstacles) if obstacles else ""
    w = ('<World type="Toric" width="%r" height="%r">%s</World>' % (float(width), float(height), inner)
         if world_type == 1 else '<World type="Infinite">%s</World>' % inner)
    pols = "\n".join(policy_xml(i, p) for i, p in enumerate(policies))
    return (f'<?xml version="1.0" encoding="utf-8"?>\n<Simulation delta_time="{dt!r}">\n  {w}\n'
            f'  <Policies>\n{pols}\n  </Policies>\n</Simulation>\n')


class RefSim:
    """One reference CrowdSimulator with agents added programmatically."""

    def __init__(self, world_type, width, height, dt, policies, agents, threads=1, xml_path=None, obstacles=None):
        xml = scenario_xml(world_type, width, height, float(np.float32(dt)), policies, obstacles)
        fd, path = tempfile.mkstemp(suffix=".xml")
        with os.fdopen(fd, "w") as f:
            f.write(xml)
        try:
            self.h = lib().ref_load(path.encode())
        finally:
            os.unlink(path)
        if not self.h:
            raise RuntimeError("reference failed to load scenario:\n" + xml)
        lib().ref_set_threads(self.h, int(threads))
        n = len(agents["id"])
        arr = lambda k, t=np.float32: np.ascontiguousarray(agents[k], dtype=t)
        self._keep = [arr("id", np.int32), arr("pos"), arr("goal"), arr("radius"), arr("pref_speed"),
                      arr("max_speed"), arr("max_accel"), arr("policy", np.int32)]
        k = self._keep
        added = lib().ref_add_agents(self.h, n, _p(k[0], _I), _p(k[1], _F), _p(k[2], _F), _p(k[3], _F),
                                     _p(k[4], _F), _p(k[5], _F), _p(k[6], _F), _p(k[7], _I))
        if added != n:
            raise RuntimeError("reference refused agent %d" % added)
        if "vel" in agents:
            vel = arr("vel")
            st = self.state()
            vd = np.ascontiguousarray(agents.get("viewdir", st["viewdir"]), dtype=np.float32)
            lib().ref_set_velocities(self.h, n, _p(vel, _F), _p(vd, _F))
        self.n = n

    @classmethod
    def from_file(cls, path, threads=1):
        """A reference simulator loaded by CrowdSimulator::FromConfigFile from `path`."""
        self = cls.__new__(cls)
        self.h = lib().ref_load(str(path).encode())
        if not self.h:
            raise RuntimeError(f"reference failed to load {path}")
        lib().ref_set_threads(self.h, int(threads))
        self._keep = []
        self.n = lib().ref_num_agents(self.h)
        return self

    def close(self):
        if self.h:
            lib().ref_destroy(self.h)
            self.h = None

    __del__ = close

    def start_csv(self, dirname, flush_immediately=False):
        lib().ref_start_csv(self.h, str(dirname).encode(), 1 if flush_immediately else 0)

    def step(self, nsteps=1):
        lib().ref_step(self.h, int(nsteps))

    def time_steps(self, nsteps):
        return lib().ref_time_steps(self.h, int(nsteps))

    def state(self):
        n = lib().ref_num_agents(self.h)
        ids = np.zeros(n, np.int32)
        out = {k: np.zeros((n, 2), np.float32) for k in ("pos", "vel", "viewdir", "pref_vel", "goal")}
        lib().ref_get_state(self.h, _p(ids, _I), _p(out["pos"], _F), _p(out["vel"], _F),
                            _p(out["viewdir"], _F), _p(out["pref_vel"], _F), _p(out["goal"], _F))
        out["id"] = ids
        return out

    def settings(self, max_policy_id=64):
        n = lib().ref_num_agents(self.h)
        out = {k: np.zeros(n, np.float32) for k in ("radius", "pref_speed", "max_speed", "max_accel")}
        out["policy_xml_id"] = np.zeros(n, np.int32)
        lib().ref_get_settings(self.h, _p(out["radius"], _F), _p(out["pref_speed"], _F), _p(out["max_speed"], _F),
                               _p(out["max_accel"], _F), _p(out["policy_xml_id"], _I), int(max_policy_id))
        return out

    def neighbors(self):
        n = lib().ref_num_agents(self.h)
        tot = lib().ref_total_neighbors(self.h)
        cnt = np.zeros(n, np.int32)
        ids = np.zeros(max(tot, 1), np.int32)
        pos = np.zeros((max(tot, 1), 2), np.float32)
        dsq = np.zeros(max(tot, 1), np.float32)
        lib().ref_get_neighbors(self.h, _p(cnt, _I), _p(ids, _I), _p(pos, _F), _p(dsq, _F))
        off = np.concatenate([[0], np.cumsum(cnt)])
        return [(ids[off[i]:off[i + 1]], pos[off[i]:off[i + 1]], dsq[off[i]:off[i + 1]]) for i in range(n)]

