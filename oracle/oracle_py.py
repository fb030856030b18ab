"""ctypes binding of the CPU oracle (oracle/libumans_oracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU-baseline leg import this module;
the product package (umans_amd) never does. The oracle restates one
WorldBase::DoStep of the reference (src/Engine/core/worldBase.cpp:79-128); see
umans_oracle.h for the precision modes.

Policies are passed in their packed numeric form (the same layout the product's
C ABI takes: method enum, sampling fields, up to four cost terms with kind,
coefficient, range and eight float parameters).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libumans_oracle.so")


class World(C.Structure):
    _fields_ = [("world_type", C.c_int32), ("width", C.c_float), ("height", C.c_float),
                ("delta_time", C.c_float), ("max_neighbors", C.c_int32), ("n_edges", C.c_int32),
                ("edges", C.POINTER(C.c_float))]


class Policy(C.Structure):
    _fields_ = [("method", C.c_int32), ("s_type", C.c_int32), ("s_base", C.c_int32),
                ("s_base_dir", C.c_int32), ("s_radius", C.c_int32), ("s_angle", C.c_float),
                ("s_speed_samples", C.c_int32), ("s_angle_samples", C.c_int32),
                ("s_random_samples", C.c_int32), ("s_include_base", C.c_int32),
                ("relaxation_time", C.c_float), ("contact_force_scale", C.c_float),
                ("n_cost", C.c_int32), ("cost_kind", C.c_int32 * 4), ("cost_coeff", C.c_float * 4),
                ("cost_range", C.c_float * 4), ("cost_params", (C.c_float * 8) * 4),
                ("gradient_iterations", C.c_int32), ("gradient_step", C.c_float)]


_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)


class AgentsIn(C.Structure):
    _fields_ = [("id", _I), ("pos", _F), ("vel", _F), ("goal", _F), ("viewdir", _F),
                ("radius", _F), ("pref_speed", _F), ("max_speed", _F), ("max_accel", _F),
                ("policy", _I), ("rng", _U)]


class AgentsOut(C.Structure):
    _fields_ = [("pos", _F), ("vel", _F), ("viewdir", _F), ("pref_vel", _F), ("next_vel", _F),
                ("contact", _F), ("rng", _U), ("vel64", _D), ("next_vel64", _D),
                ("list_stride", C.c_int32), ("nb_count", _I), ("nb_id", _I), ("nb_d2", _D),
                ("nb_seg", _I), ("nb_full", _I)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.oracle_step.argtypes = [C.POINTER(World), C.POINTER(Policy), C.c_int, C.c_int,
                                     C.POINTER(AgentsIn), C.POINTER(AgentsOut), C.c_int, C.c_int]
        _lib.oracle_step.restype = C.c_int
        _lib.oracle_neighbors_bruteforce.argtypes = [C.POINTER(World), C.POINTER(Policy), C.c_int,
                                                     C.POINTER(AgentsIn), C.c_int, _I, _I, _D, _I]
        _lib.oracle_neighbors_bruteforce.restype = C.c_int
        _lib.oracle_rng_seed.argtypes = [C.c_uint32]
        _lib.oracle_rng_seed.restype = C.c_uint32
        _lib.oracle_rng_uniform.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_int, _F]
        _lib.oracle_rng_uniform.restype = C.c_uint32
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def make_world(world_type, width, height, dt, max_neighbors, edges=None):
    """edges: optional (E, 4) float32 obstacle edges (ax, ay, bx, by) in polygon order."""
    w = World(int(world_type), float(width), float(height), float(dt), int(max_neighbors), 0, None)
    if edges is not None and len(edges):
        e = np.ascontiguousarray(edges, np.float32).reshape(-1, 4)
        w._edges = e
        w.n_edges = len(e)
        w.edges = e.ctypes.data_as(C.POINTER(C.c_float))
    return w


def polygon_edges(counts, points):
    """Polygon2D edges (Polygon2D.cpp:27-38): vertex i to vertex (i + 1) mod n, per polygon."""
    out, o = [], 0
    pts = np.asarray(points, np.float32).reshape(-1, 2)
    for c in counts:
        for i in range(c):
            a, b = pts[o + i], pts[o + (i + 1) % c]
            out.append([a[0], a[1], b[0], b[1]])
        o += c
    return np.asarray(out, np.float32).reshape(-1, 4)


def make_policies(packed):
    """packed: list of dicts with the numeric policy fields (see Policy._fields_)."""
    arr = (Policy * len(packed))()
    for i, p in enumerate(packed):
        q = arr[i]
        for k in ("method", "s_type", "s_base", "s_base_dir", "s_radius", "s_speed_samples",
                  "s_angle_samples", "s_random_samples", "s_include_base", "n_cost"):
            setattr(q, k, int(p[k]))
        for k in ("s_angle", "relaxation_time", "contact_force_scale"):
            setattr(q, k, float(p[k]))
        q.gradient_iterations = int(p.get("gradient_iterations", 1))
        q.gradient_step = float(p.get("gradient_step", 1.0))
        for c in range(p["n_cost"]):
            q.cost_kind[c] = int(p["cost_kind"][c])
            q.cost_coeff[c] = float(p["cost_coeff"][c])
            q.cost_range[c] = float(p["cost_range"][c])
            for j in range(8):
                q.cost_params[c][j] = float(p["cost_params"][c][j])
    return arr


def _agents_in(a):
    keep = {}
    def f32(k, shape2=False):
        x = np.ascontiguousarray(a[k], dtype=np.float32)
        keep[k] = x
        return _ptr(x, _F)
    ids = np.ascontiguousarray(a["id"], dtype=np.int32); keep["id"] = ids
    pol = np.ascontiguousarray(a["policy"], dtype=np.int32); keep["policy"] = pol
    rng = a.get("rng")
    if rng is not None:
        rng = np.ascontiguousarray(rng, dtype=np.uint32); keep["rng"] = rng
    viewdir = a.get("viewdir")
    if viewdir is None:
        viewdir = np.zeros_like(np.asarray(a["pos"], dtype=np.float32))
        viewdir[:, 0] = 1.0
    keep["viewdir"] = np.ascontiguousarray(viewdir, dtype=np.float32)
    s = AgentsIn(_ptr(ids, _I), f32("pos"), f32("vel"), f32("goal"),
                 _ptr(keep["viewdir"], _F), f32("radius"), f32("pref_speed"), f32("max_speed"),
                 f32("max_accel"), _ptr(pol, _I), _ptr(rng, _U))
    return s, keep


def step(world, policies, agents, precision=0, nthreads=0, list_stride=0):
    """One oracle DoStep. agents: dict of numpy arrays (id, pos[n,2], vel, goal, viewdir,
    radius, pref_speed, max_speed, max_accel, policy, optional rng).
    Returns a dict of outputs; with list_stride > 0 also the neighbour lists."""
    n = len(agents["id"])
    ain, keep = _agents_in(agents)
    out = {k: np.zeros((n, 2), np.float32) for k in ("pos", "vel", "viewdir", "pref_vel", "next_vel", "contact")}
    out["vel64"] = np.zeros((n, 2), np.float64)
    out["next_vel64"] = np.zeros((n, 2), np.float64)
    out["rng"] = np.zeros(n, np.uint32)
    if list_stride > 0:
        out["nb_count"] = np.zeros(n, np.int32)
        out["nb_id"] = np.full((n, list_stride), -1, np.int32)
        out["nb_d2"] = np.zeros((n, list_stride), np.float64)
        out["nb_seg"] = np.zeros((n, list_stride), np.int32)
        out["nb_full"] = np.zeros(n, np.int32)
    g = lambda k, t: _ptr(out.get(k), t)
    ao = AgentsOut(g("pos", _F), g("vel", _F), g("viewdir", _F), g("pref_vel", _F), g("next_vel", _F),
                   g("contact", _F), g("rng", _U) if agents.get("rng") is not None else None,
                   g("vel64", _D), g("next_vel64", _D), int(list_stride),
                   g("nb_count", _I), g("nb_id", _I), g("nb_d2", _D), g("nb_seg", _I), g("nb_full", _I))
    pols = make_policies(policies) if not isinstance(policies, C.Array) else policies
    rc = lib().oracle_step(C.byref(world), pols, len(pols), n, C.byref(ain), C.byref(ao),
                           int(precision), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle_step failed: {rc}")
    return out


def neighbors_bruteforce(world, policies, agents, list_stride):
    n = len(agents["id"])
    ain, keep = _agents_in(agents)
    cnt = np.zeros(n, np.int32)
    ids = np.full((n, list_stride), -1, np.int32)
    d2 = np.zeros((n, list_stride), np.float64)
    seg = np.zeros((n, list_stride), np.int32)
    pols = make_policies(policies)
    lib().oracle_neighbors_bruteforce(C.byref(world), pols, n, C.byref(ain), list_stride,
                                      _ptr(cnt, _I), _ptr(ids, _I), _ptr(d2, _D), _ptr(seg, _I))
    return cnt, ids, d2, seg


def rng_seed(agent_id):
    return lib().oracle_rng_seed(int(agent_id))


def rng_uniform(state, lo, hi, count):
    out = np.zeros(count, np.float32)
    st = lib().oracle_rng_uniform(int(state), float(lo), float(hi), int(count), _ptr(out, _F))
    return st, out
