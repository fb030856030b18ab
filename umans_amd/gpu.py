"""ctypes binding of libumans_gpu.so (the C ABI in include/umans_gpu.h).

`World` mirrors what the reference's CrowdSimulator + WorldBase expose for the per-step
loop (crowdSimulator.h:58-96, worldBase.h:191-229): policies by registry name and
parameters, agents added in bulk, DoStep, and state/neighbour read-back. Every call
goes through the HIP library; if the library is missing this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# UMANS_GPU_LIB names another build of the same library (for example one compiled with
# different launch constants); the default is the in-tree build.
LIB_PATH = os.environ.get("UMANS_GPU_LIB") or os.path.join(HERE, "libumans_gpu.so")

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)

EXPORTS = [
    "umans_create", "umans_destroy", "umans_last_error", "umans_set_policies", "umans_upload",
    "umans_download", "umans_num_agents", "umans_step", "umans_synchronize", "umans_get_stats",
    "umans_set_record_neighbors", "umans_dump_neighbors", "umans_cost_id", "umans_cost_name",
    "umans_cost_param_index", "umans_cost_param_name", "umans_cost_defaults", "umans_policy_init",
    "umans_set_profiling", "umans_get_profile", "umans_create_batch", "umans_num_worlds",
    "umans_set_graph", "umans_autotune", "umans_set_tuning", "umans_adjoint_forward", "umans_adjoint_backward",
    "umans_cost_param_read_name", "umans_scenario_load", "umans_scenario_free", "umans_scenario_get_info",
    "umans_scenario_warning", "umans_scenario_get", "umans_scenario_obstacles", "umans_csv_start",
    "umans_csv_flush", "umans_csv_stop", "umans_add_agents", "umans_remove_agent", "umans_set_agent_goal",
    "umans_set_agent_states", "umans_set_remove_at_goal", "umans_num_scheduled", "umans_download_records",
    "umans_set_obstacles", "umans_step_noalloc", "umans_adjoint_clamps",
]

ENOTFOUND = -2


class WorldDesc(C.Structure):
    _fields_ = [("world_type", C.c_int32), ("width", C.c_float), ("height", C.c_float),
                ("delta_time", C.c_float), ("max_neighbors", C.c_int32), ("cell_size", C.c_float),
                ("block_size", C.c_int32), ("device", C.c_int32)]


class PolicyC(C.Structure):
    _fields_ = [("method", C.c_int32), ("s_type", C.c_int32), ("s_base", C.c_int32),
                ("s_base_dir", C.c_int32), ("s_radius", C.c_int32), ("s_angle", C.c_float),
                ("s_speed_samples", C.c_int32), ("s_angle_samples", C.c_int32),
                ("s_random_samples", C.c_int32), ("s_include_base", C.c_int32),
                ("relaxation_time", C.c_float), ("contact_force_scale", C.c_float),
                ("n_cost", C.c_int32), ("cost_kind", C.c_int32 * 4), ("cost_coeff", C.c_float * 4),
                ("cost_range", C.c_float * 4), ("cost_params", (C.c_float * 8) * 4),
                ("gradient_iterations", C.c_int32), ("gradient_step", C.c_float)]


class AgentsC(C.Structure):
    _fields_ = [("id", _I), ("pos", _F), ("vel", _F), ("goal", _F), ("viewdir", _F), ("radius", _F),
                ("pref_speed", _F), ("max_speed", _F), ("max_accel", _F), ("policy", _I), ("rng", _U),
                ("pref_vel", _F), ("world", _I)]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("overflow_agents", C.c_int64), ("max_list", C.c_int64),
                ("sim_time", C.c_double), ("obstacle_overflow_agents", C.c_int64)]


class Profile(C.Structure):
    _fields_ = [("steps", C.c_int64), ("grid_ms", C.c_double), ("step_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("cell_factor", C.c_float), ("block_size", C.c_int32), ("ms_per_step", C.c_float),
                ("list_capacity", C.c_int32), ("class_block", C.c_int32 * 5)]


class AdjointDesc(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("checkpoint_every", C.c_int32), ("param_kind", C.c_int32)]


class ScenarioInfo(C.Structure):
    _fields_ = [("world_type", C.c_int32), ("width", C.c_float), ("height", C.c_float), ("delta_time", C.c_float),
                ("end_time", C.c_float), ("n_agents", C.c_int32), ("n_policies", C.c_int32),
                ("n_obstacles", C.c_int32), ("n_warnings", C.c_int32)]


class UmansError(RuntimeError):
    pass


class UmansTruncated(UmansError):
    """UMANS_ETRUNC: a step kept fewer neighbours than the reference would (an uncapped
    list longer than 4096 entries, or a list longer than the adjoint keeps). The call's
    outputs are still written."""


ETRUNC = -1002


_lib = None


def lib():
    """Load libumans_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UmansError(f"{LIB_PATH} is missing: build it with `python -m umans_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.umans_create.argtypes = [C.POINTER(WorldDesc), C.POINTER(vp)]
        L.umans_destroy.argtypes = [vp]
        L.umans_destroy.restype = None
        L.umans_last_error.argtypes = [vp]
        L.umans_last_error.restype = C.c_char_p
        L.umans_set_policies.argtypes = [vp, C.POINTER(PolicyC), C.c_int]
        L.umans_upload.argtypes = [vp, C.POINTER(AgentsC), C.c_int]
        L.umans_download.argtypes = [vp, C.POINTER(AgentsC)]
        L.umans_num_agents.argtypes = [vp]
        L.umans_step.argtypes = [vp, C.c_int]
        L.umans_synchronize.argtypes = [vp]
        L.umans_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.umans_set_record_neighbors.argtypes = [vp, C.c_int]
        L.umans_dump_neighbors.argtypes = [vp, C.c_int, _I, _I, _D, _I]
        L.umans_cost_id.argtypes = [C.c_char_p]
        L.umans_cost_name.argtypes = [C.c_int]
        L.umans_cost_name.restype = C.c_char_p
        L.umans_cost_param_index.argtypes = [C.c_int, C.c_char_p]
        L.umans_cost_param_name.argtypes = [C.c_int, C.c_int]
        L.umans_cost_param_name.restype = C.c_char_p
        L.umans_cost_defaults.argtypes = [C.c_int, _F, _F]
        L.umans_policy_init.argtypes = [C.POINTER(PolicyC), C.c_int]
        L.umans_create_batch.argtypes = [C.POINTER(WorldDesc), C.c_int, C.POINTER(vp)]
        L.umans_num_worlds.argtypes = [vp]
        L.umans_set_graph.argtypes = [vp, C.c_int]
        L.umans_autotune.argtypes = [vp, C.c_int, C.POINTER(Tuning)]
        L.umans_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.umans_adjoint_forward.argtypes = [vp, C.POINTER(AdjointDesc), _F, C.POINTER(C.c_double)]
        L.umans_adjoint_backward.argtypes = [vp, _F, _F, _F]
        L.umans_adjoint_clamps.argtypes = [vp, C.POINTER(C.c_uint8)]
        L.umans_cost_param_read_name.argtypes = [C.c_int, C.c_int]
        L.umans_cost_param_read_name.restype = C.c_char_p
        L.umans_scenario_load.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int]
        L.umans_scenario_free.argtypes = [vp]
        L.umans_scenario_free.restype = None
        L.umans_scenario_get_info.argtypes = [vp, C.POINTER(ScenarioInfo)]
        L.umans_scenario_warning.argtypes = [vp, C.c_int]
        L.umans_scenario_warning.restype = C.c_char_p
        L.umans_scenario_get.argtypes = [vp, C.POINTER(PolicyC), _I, C.POINTER(AgentsC), _F, _I]
        L.umans_scenario_obstacles.argtypes = [vp, _I, _F]
        L.umans_add_agents.argtypes = [vp, C.POINTER(AgentsC), C.c_int, _F, C.POINTER(C.c_uint8), C.c_int, _I]
        L.umans_remove_agent.argtypes = [vp, C.c_int32]
        L.umans_set_agent_goal.argtypes = [vp, C.c_int32, C.c_float, C.c_float]
        L.umans_set_agent_states.argtypes = [vp, C.c_int, _I, _F, _F, _F]
        L.umans_set_remove_at_goal.argtypes = [vp, C.c_int, _I, C.POINTER(C.c_uint8)]
        L.umans_num_scheduled.argtypes = [vp]
        L.umans_download_records.argtypes = [vp, vp, C.c_int]
        L.umans_set_obstacles.argtypes = [vp, C.c_int, _I, _F]
        L.umans_csv_start.argtypes = [vp, C.c_char_p, C.c_int]
        L.umans_csv_flush.argtypes = [vp]
        L.umans_csv_stop.argtypes = [vp]
        L.umans_set_profiling.argtypes = [vp, C.c_int]
        L.umans_get_profile.argtypes = [vp, C.POINTER(Profile)]
        L.umans_step_noalloc.argtypes = [C.POINTER(WorldDesc), C.POINTER(PolicyC), C.c_int, C.POINTER(AgentsC),
                                         C.c_int, C.c_int, vp, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def policy_struct(packed):
    q = PolicyC()
    for k in ("method", "s_type", "s_base", "s_base_dir", "s_radius", "s_speed_samples",
              "s_angle_samples", "s_random_samples", "s_include_base", "n_cost"):
        setattr(q, k, int(packed[k]))
    for k in ("s_angle", "relaxation_time", "contact_force_scale"):
        setattr(q, k, float(packed[k]))
    q.gradient_iterations = int(packed.get("gradient_iterations", 1))
    q.gradient_step = float(packed.get("gradient_step", 1.0))
    for c in range(4):
        q.cost_kind[c] = int(packed["cost_kind"][c])
        q.cost_coeff[c] = float(packed["cost_coeff"][c])
        q.cost_range[c] = float(packed["cost_range"][c])
        for j in range(8):
            q.cost_params[c][j] = float(packed["cost_params"][c][j])
    return q


AGENT_FIELDS = ("id", "pos", "vel", "goal", "viewdir", "radius", "pref_speed", "max_speed", "max_accel", "policy",
                "rng", "pref_vel", "world")


def step_noalloc(world_type, width, height, dt, policies, max_neighbors, ptrs, n, nsteps=1, stream=0,
                 cell_size=0.0, block_size=0):
    """umans_step_noalloc: nsteps steps of one world on caller-owned device arrays.
    `ptrs` maps umans_agents field names to device addresses (ints; a missing field is
    null); `stream` is a hipStream_t as an int (0: the null stream). Asynchronous."""
    L = lib()
    desc = WorldDesc(int(world_type), float(width), float(height), float(dt), int(max_neighbors), float(cell_size),
                     int(block_size), 0)
    types = {"id": _I, "policy": _I, "world": _I, "rng": _U}
    a = AgentsC(*[C.cast(C.c_void_p(ptrs[f]), types.get(f, _F)) if ptrs.get(f) else None for f in AGENT_FIELDS])
    pols = (PolicyC * len(policies))(*[policy_struct(p) for p in policies])
    err = C.create_string_buffer(256)
    rc = L.umans_step_noalloc(C.byref(desc), pols, len(policies), C.byref(a), int(n), int(nsteps),
                              C.c_void_p(stream or None), err, len(err))
    if rc != 0:
        raise UmansError(f"umans_step_noalloc failed ({rc}): {err.value.decode()}")


class World:
    """One simulated world on one GPU (a context of the C ABI)."""

    def __init__(self, world_type, width, height, dt, policies, max_neighbors=0, cell_size=0.0, device=0,
                 block_size=0, batch=None):
        """batch: optional list of (world_type, width, height, dt) for a batch of worlds
        (then world_type/width/height/dt are ignored and agents carry a `world` index)."""
        L = lib()
        self._h = C.c_void_p()
        if batch is None:
            batch = [(world_type, width, height, dt)]
        descs = (WorldDesc * len(batch))(*[WorldDesc(int(t), float(w), float(h), float(d), int(max_neighbors),
                                                     float(cell_size), int(block_size), int(device))
                                           for (t, w, h, d) in batch])
        rc = L.umans_create_batch(descs, len(batch), C.byref(self._h))
        if rc != 0:
            raise UmansError(f"umans_create_batch failed ({rc})")
        self.desc = descs[0]
        self.n_worlds = len(batch)
        self._check(L.umans_set_policies(self._h, (PolicyC * len(policies))(*[policy_struct(p) for p in policies]),
                                         len(policies)))

    def _check(self, rc):
        if rc != 0:
            msg = lib().umans_last_error(self._h)
            cls = UmansTruncated if rc == ETRUNC else UmansError
            raise cls(f"umans call failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            lib().umans_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        """Agents in the list (WorldBase::GetAgents), scheduled ones not counted."""
        return int(lib().umans_num_agents(self._h))

    @classmethod
    def from_scenario_file(cls, path, max_neighbors=0, device=0):
        """A world loaded from a scenario file of the reference's XML schema, in the
        reference's load order: agents whose start time is 0 join the list in file
        order, later ones are scheduled (CrowdSimulator::FromConfigFile ->
        WorldBase::AddAgent, crowdSimulator.cpp:417-512); remove_at_goal is kept per
        agent; obstacles are set on the world."""
        sc, meta = load_scenario(path)
        w = cls(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, max_neighbors, device=device)
        if len(meta["obstacle_counts"]):
            w.set_obstacles(meta["obstacle_counts"], meta["obstacle_points"])
        later = np.asarray(meta["start_time"], np.float32) > 0
        rm = np.asarray(meta["remove_at_goal"]).astype(bool)
        now = {k: (np.asarray(v)[~later] if v is not None else None) for k, v in sc.agents.items()}
        w.upload(now)
        if rm[~later].any():
            w.set_remove_at_goal(now["id"], rm[~later])
        if later.any():
            sub = {k: (np.asarray(v)[later] if v is not None else None) for k, v in sc.agents.items()}
            w.add_agents(sub, start_time=np.asarray(meta["start_time"], np.float32)[later], remove_at_goal=rm[later],
                         keep_ids=True)
        w.scenario = sc
        w.scenario_meta = meta
        return w

    def add_agents(self, agents, start_time=None, remove_at_goal=None, keep_ids=False):
        """WorldBase::AddAgent for each agent (dict of arrays; "id" optional = desired
        IDs, "goal" optional = the position). keep_ids: the IDs are final (assigned by
        the scenario loader). Returns the IDs given."""
        n = len(agents["pos"])
        keep = {}

        def arr(k, t, dtype):
            if k not in agents or agents[k] is None:
                return None
            keep[k] = np.ascontiguousarray(agents[k], dtype=dtype)
            return _p(keep[k], t)
        a = AgentsC(arr("id", _I, np.int32), arr("pos", _F, np.float32), arr("vel", _F, np.float32),
                    arr("goal", _F, np.float32), arr("viewdir", _F, np.float32), arr("radius", _F, np.float32),
                    arr("pref_speed", _F, np.float32), arr("max_speed", _F, np.float32),
                    arr("max_accel", _F, np.float32), arr("policy", _I, np.int32), arr("rng", _U, np.uint32),
                    None, None)
        st = None if start_time is None else np.ascontiguousarray(start_time, np.float32)
        rm = None if remove_at_goal is None else np.ascontiguousarray(remove_at_goal, np.uint8)
        ids = np.zeros(n, np.int32)
        self._check(lib().umans_add_agents(self._h, C.byref(a), n, None if st is None else _p(st, _F),
                                           None if rm is None else rm.ctypes.data_as(C.POINTER(C.c_uint8)),
                                           1 if keep_ids else 0, _p(ids, _I)))
        return ids

    def remove_agent(self, agent_id):
        """WorldBase::RemoveAgent: False when no agent has that ID."""
        rc = lib().umans_remove_agent(self._h, int(agent_id))
        if rc == ENOTFOUND:
            return False
        self._check(rc)
        return True

    def set_agent_goal(self, agent_id, x, y):
        rc = lib().umans_set_agent_goal(self._h, int(agent_id), float(x), float(y))
        if rc == ENOTFOUND:
            return False
        self._check(rc)
        return True

    def set_agent_states(self, ids, pos, vel, viewdir):
        ids = np.ascontiguousarray(ids, np.int32)
        pos, vel, viewdir = (np.ascontiguousarray(x, np.float32) for x in (pos, vel, viewdir))
        self._check(lib().umans_set_agent_states(self._h, len(ids), _p(ids, _I), _p(pos, _F), _p(vel, _F),
                                                 _p(viewdir, _F)))

    def set_remove_at_goal(self, ids, flags):
        ids = np.ascontiguousarray(ids, np.int32)
        f = np.ascontiguousarray(flags, np.uint8)
        self._check(lib().umans_set_remove_at_goal(self._h, len(ids), _p(ids, _I),
                                                   f.ctypes.data_as(C.POINTER(C.c_uint8))))

    def records(self):
        """The list as (n, 7) float32 view of packed records: id (int32 bits), position,
        velocity, viewing direction."""
        n = self.n
        buf = np.zeros((n, 7), np.float32)
        self._check(lib().umans_download_records(self._h, buf.ctypes.data_as(C.c_void_p), n))
        return buf[:, 0].view(np.int32).copy(), buf[:, 1:3].copy(), buf[:, 3:5].copy(), buf[:, 5:7].copy()

    def set_obstacles(self, counts, points):
        """Polygons as vertex counts and (sum(counts), 2) vertices; edges i -> (i + 1) mod n."""
        cnt = np.ascontiguousarray(counts, np.int32)
        pts = np.ascontiguousarray(points, np.float32).reshape(-1)
        self._check(lib().umans_set_obstacles(self._h, len(cnt), _p(cnt, _I), _p(pts, _F)))

    def num_scheduled(self):
        return int(lib().umans_num_scheduled(self._h))

    def upload(self, agents):
        n = len(agents["id"])
        keep = {}

        def arr(k, t, dtype, required=True):
            if k not in agents or agents[k] is None:
                if required:
                    raise KeyError(k)
                return None
            keep[k] = np.ascontiguousarray(agents[k], dtype=dtype)
            return _p(keep[k], t)
        a = AgentsC(arr("id", _I, np.int32), arr("pos", _F, np.float32), arr("vel", _F, np.float32, False),
                    arr("goal", _F, np.float32), arr("viewdir", _F, np.float32, False),
                    arr("radius", _F, np.float32), arr("pref_speed", _F, np.float32),
                    arr("max_speed", _F, np.float32), arr("max_accel", _F, np.float32),
                    arr("policy", _I, np.int32), arr("rng", _U, np.uint32, False), None,
                    arr("world", _I, np.int32, False))
        self._check(lib().umans_upload(self._h, C.byref(a), n))

    def step(self, nsteps=1):
        self._check(lib().umans_step(self._h, int(nsteps)))

    def synchronize(self):
        self._check(lib().umans_synchronize(self._h))

    def record_neighbors(self, enable=True):
        self._check(lib().umans_set_record_neighbors(self._h, 1 if enable else 0))

    def download(self):
        n = self.n
        out = {k: np.zeros((n, 2), np.float32) for k in ("pos", "vel", "viewdir", "goal", "pref_vel")}
        for k in ("radius", "pref_speed", "max_speed", "max_accel"):
            out[k] = np.zeros(n, np.float32)
        out["id"] = np.zeros(n, np.int32)
        out["policy"] = np.zeros(n, np.int32)
        out["rng"] = np.zeros(n, np.uint32)
        out["world"] = np.zeros(n, np.int32)
        a = AgentsC(_p(out["id"], _I), _p(out["pos"], _F), _p(out["vel"], _F), _p(out["goal"], _F),
                    _p(out["viewdir"], _F), _p(out["radius"], _F), _p(out["pref_speed"], _F),
                    _p(out["max_speed"], _F), _p(out["max_accel"], _F), _p(out["policy"], _I),
                    _p(out["rng"], _U), _p(out["pref_vel"], _F), _p(out["world"], _I))
        self._check(lib().umans_download(self._h, C.byref(a)))
        return out

    def neighbors(self, stride=128):
        n = self.n
        cnt = np.zeros(n, np.int32)
        ids = np.full((n, stride), -1, np.int32)
        d2 = np.zeros((n, stride), np.float64)
        seg = np.zeros((n, stride), np.int32)
        self._check(lib().umans_dump_neighbors(self._h, stride, _p(cnt, _I), _p(ids, _I), _p(d2, _D), _p(seg, _I)))
        return cnt, ids, d2, seg

    def csv_start(self, dirname, flush_immediately=False):
        self._check(lib().umans_csv_start(self._h, str(dirname).encode(), 1 if flush_immediately else 0))

    def csv_flush(self):
        self._check(lib().umans_csv_flush(self._h))

    def csv_stop(self):
        self._check(lib().umans_csv_stop(self._h))

    def set_graph(self, enable=True):
        self._check(lib().umans_set_graph(self._h, 1 if enable else 0))

    def autotune(self, trial_steps=3):
        t = Tuning()
        self._check(lib().umans_autotune(self._h, int(trial_steps), C.byref(t)))
        return {"cell_factor": t.cell_factor, "block_size": t.block_size, "ms_per_step": t.ms_per_step,
                "list_capacity": t.list_capacity, "class_block": list(t.class_block)}

    def set_tuning(self, cell_factor, block_size=0, list_capacity=0, class_block=(0, 0, 0, 0, 0)):
        """class_block: block size per policy class (force + gradient, force + sampling,
        ORCA, generic, cooperative sampling); 0 = block_size."""
        t = Tuning(float(cell_factor), int(block_size), 0.0, int(list_capacity),
                   (C.c_int32 * 5)(*[int(b) for b in class_block]))
        self._check(lib().umans_set_tuning(self._h, C.byref(t)))

    def set_policies(self, policies):
        self._check(lib().umans_set_policies(self._h, (PolicyC * len(policies))(*[policy_struct(p) for p in policies]),
                                             len(policies)))

    def adjoint_forward(self, targets, nsteps, checkpoint_every=10, param_kind=2):
        """Run nsteps from the current state with checkpoints; return the position loss
        against targets (nsteps x n x 2, upload order). param_kind: the cost whose two
        parameters are differentiated (2 SocialForcesAvoidance, 1 PowerLaw)."""
        t = np.ascontiguousarray(targets, dtype=np.float32)
        if t.shape != (nsteps, self.n, 2):
            raise ValueError(f"targets must have shape {(nsteps, self.n, 2)}")
        self._adj_keep = t
        return self.adjoint_forward_ptr(t.ctypes.data, nsteps, checkpoint_every, param_kind)

    def adjoint_forward_ptr(self, targets_ptr, nsteps, checkpoint_every=10, param_kind=2):
        """adjoint_forward with the targets given by address (host or device memory of
        nsteps x n x 2 floats, ready when called)."""
        loss = C.c_double()
        d = AdjointDesc(int(nsteps), int(checkpoint_every), int(param_kind))
        self._check(lib().umans_adjoint_forward(self._h, C.byref(d), C.cast(C.c_void_p(targets_ptr), _F),
                                                C.byref(loss)))
        return loss.value

    def adjoint_backward(self, n_policies):
        dp = np.zeros((n_policies, 2), np.float32)
        dpos = np.zeros((self.n, 2), np.float32)
        dvel = np.zeros((self.n, 2), np.float32)
        self._check(lib().umans_adjoint_backward(self._h, _p(dp, _F), _p(dpos, _F), _p(dvel, _F)))
        return dp, dpos, dvel

    def adjoint_clamps(self, nsteps):
        """The clamp branches of the last backward pass: (nsteps, n) uint8, bit 0 the
        acceleration clamp, bit 1 the speed clamp."""
        out = np.zeros((nsteps, self.n), np.uint8)
        self._check(lib().umans_adjoint_clamps(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def adjoint_backward_ptr(self, n_policies, dpos_ptr, dvel_ptr):
        """adjoint_backward writing dL/dpos0 and dL/dvel0 to the given addresses (host or
        device memory of n x 2 floats; 0 = not wanted); returns dL/d(parameters)."""
        dp = np.zeros((n_policies, 2), np.float32)
        cast = lambda q: C.cast(C.c_void_p(q), _F) if q else None
        self._check(lib().umans_adjoint_backward(self._h, _p(dp, _F), cast(dpos_ptr), cast(dvel_ptr)))
        return dp

    def set_profiling(self, enable=True):
        self._check(lib().umans_set_profiling(self._h, 1 if enable else 0))

    def profile(self):
        p = Profile()
        self._check(lib().umans_get_profile(self._h, C.byref(p)))
        return {"steps": p.steps, "grid_ms": p.grid_ms, "step_ms": p.step_ms}

    def stats(self):
        s = Stats()
        self._check(lib().umans_get_stats(self._h, C.byref(s)))
        return {"steps": s.steps, "overflow_agents": s.overflow_agents, "max_list": s.max_list,
                "sim_time": s.sim_time, "obstacle_overflow_agents": s.obstacle_overflow_agents}


def world_from_scenario(sc, cell_size=0.0, device=0):
    w = World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors, cell_size, device)
    w.upload(sc.agents)
    return w


def packed_from_struct(q):
    """Packed-policy dict (umans_amd.registry layout) from a umans_policy struct."""
    d = {k: getattr(q, k) for k in ("method", "s_type", "s_base", "s_base_dir", "s_radius", "s_angle",
                                    "s_speed_samples", "s_angle_samples", "s_random_samples", "s_include_base",
                                    "relaxation_time", "contact_force_scale", "n_cost",
                                    "gradient_iterations", "gradient_step")}
    d["cost_kind"] = list(q.cost_kind)
    d["cost_coeff"] = list(q.cost_coeff)
    d["cost_range"] = list(q.cost_range)
    d["cost_params"] = [list(q.cost_params[c]) for c in range(4)]
    return d


def load_scenario(path):
    """Read a scenario file of the reference's XML schema (CrowdSimulator::FromConfigFile)
    with the library's host loader. Returns (Scenario, info dict with end_time, policy XML
    ids, obstacles, warnings, start_time and remove_at_goal per agent)."""
    from .scenarios import Scenario
    L = lib()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.umans_scenario_load(str(path).encode(), C.byref(h), err, 512)
    if rc != 0:
        raise UmansError(f"cannot load {path}: {err.value.decode()}")
    try:
        info = ScenarioInfo()
        L.umans_scenario_get_info(h, C.byref(info))
        n, npol = info.n_agents, info.n_policies
        pols = (PolicyC * max(npol, 1))()
        xml_ids = np.zeros(max(npol, 1), np.int32)
        a = {"id": np.zeros(n, np.int32), "pos": np.zeros((n, 2), np.float32), "vel": np.zeros((n, 2), np.float32),
             "goal": np.zeros((n, 2), np.float32), "viewdir": np.zeros((n, 2), np.float32),
             "radius": np.zeros(n, np.float32), "pref_speed": np.zeros(n, np.float32),
             "max_speed": np.zeros(n, np.float32), "max_accel": np.zeros(n, np.float32),
             "policy": np.zeros(n, np.int32), "rng": np.zeros(n, np.uint32)}
        st = np.zeros(n, np.float32)
        rag = np.zeros(n, np.int32)
        ac = AgentsC(_p(a["id"], _I), _p(a["pos"], _F), _p(a["vel"], _F), _p(a["goal"], _F), _p(a["viewdir"], _F),
                     _p(a["radius"], _F), _p(a["pref_speed"], _F), _p(a["max_speed"], _F), _p(a["max_accel"], _F),
                     _p(a["policy"], _I), _p(a["rng"], _U), None, None)
        L.umans_scenario_get(h, pols, _p(xml_ids, _I), C.byref(ac), _p(st, _F), _p(rag, _I))
        counts = np.zeros(max(info.n_obstacles, 1), np.int32)
        L.umans_scenario_obstacles(h, _p(counts, _I), None)
        pts = np.zeros((int(counts[:info.n_obstacles].sum()), 2), np.float32)
        L.umans_scenario_obstacles(h, None, _p(pts, _F))
        warnings = [L.umans_scenario_warning(h, i).decode() for i in range(info.n_warnings)]
    finally:
        L.umans_scenario_free(h)
    packed = [packed_from_struct(pols[i]) for i in range(npol)]
    sc = Scenario(str(path), info.world_type, info.width, info.height, info.delta_time, 0, packed, a)
    meta = {"end_time": info.end_time, "policy_xml_ids": xml_ids[:npol].tolist(), "start_time": st,
            "remove_at_goal": rag, "obstacle_counts": counts[:info.n_obstacles], "obstacle_points": pts,
            "warnings": warnings}
    sc.meta.update(meta)
    return sc, meta
