"""Seeded synthetic crowds in the reference's terms (agents + policies + world).

Each generator returns a Scenario: per-agent numpy arrays, the packed policies
(umans_amd.registry.pack_policy) and the world description. Seeds are fixed per
workload so every machine regenerates the same inputs; no data files are needed.
Agent defaults are the reference's Agent::Settings (radius 0.24, preferred speed 1.4,
maximum speed 1.8, maximum acceleration 5; src/Engine/core/agent.h:39-57).
"""
from dataclasses import dataclass, field

import numpy as np

from .registry import pack_policy


@dataclass
class Scenario:
    name: str
    world_type: int              # 0 infinite, 1 toric
    width: float
    height: float
    dt: float
    max_neighbors: int
    policies: list               # packed policies
    agents: dict                 # id, pos, vel, goal, viewdir, radius, pref_speed, max_speed, max_accel, policy
    steps: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.agents["id"])


def _agents(n):
    return {
        "id": np.arange(n, dtype=np.int32),
        "pos": np.zeros((n, 2), np.float32),
        "vel": np.zeros((n, 2), np.float32),
        "goal": np.zeros((n, 2), np.float32),
        "viewdir": np.zeros((n, 2), np.float32),
        "radius": np.full(n, 0.24, np.float32),
        "pref_speed": np.full(n, 1.4, np.float32),
        "max_speed": np.full(n, 1.8, np.float32),
        "max_accel": np.full(n, 5.0, np.float32),
        "policy": np.zeros(n, np.int32),
    }


def initial_viewdir(pos, goal):
    """Agent::setGoal looks straight at the goal (agent.cpp:120-127): float normalisation."""
    d = (goal - pos).astype(np.float32)
    m = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)).astype(np.float32)
    out = np.zeros_like(d)
    nz = m > 0
    out[nz, 0] = d[nz, 0] / m[nz]
    out[nz, 1] = d[nz, 1] / m[nz]
    return out.astype(np.float32)


# --------------------------------------------------------------- policies
def orca_policy(time_horizon=5.0, rng=8.0):
    return pack_policy("global", [("ORCA", 1.0, {"timeHorizon": time_horizon, "range": rng})])


def powerlaw_sampling_policy(k=1.5, tau0=3.0, rng=4.0, angles=12, speeds=6, relax=0.5):
    """{PowerLaw + GoalReachingForce} with regular sampling over the full circle of
    maximum-speed radius, base zero, direction = preferred velocity."""
    return pack_policy("sampling",
                       [("PowerLaw", 1.0, {"k": k, "tau0": tau0, "range": rng}),
                        ("GoalReachingForce", 1.0, {})],
                       sampling={"SamplingType": "regular", "SamplingBase": "zero",
                                 "SamplingBaseDirection": "preferred velocity", "SamplingAngle": 360,
                                 "SpeedSamples": speeds, "AngleSamples": angles,
                                 "SamplingRadius": "maximum speed"},
                       relaxation_time=relax)


def socialforces_gradient_policy(V0=2000.0, sigma=0.08, rng=4.0, mass=80.0, relax=0.5):
    """{SocialForcesAvoidance + GoalReachingForce} with the gradient method. A and B of the
    Helbing model map to V0 and sigma; agentMass 80 gives forces in newtons."""
    return pack_policy("gradient",
                       [("SocialForcesAvoidance", 1.0, {"V0": V0, "sigma": sigma, "range": rng,
                                                       "agentMass": mass}),
                        ("GoalReachingForce", 1.0, {})],
                       relaxation_time=relax)


def karamouzas_sampling_policy(rng=5.0, angles=16, speeds=8):
    """Karamouzas.xml geometry (base zero, direction = preferred velocity, 180 degrees,
    radius = maximum speed, base included) with 16 angles x 8 speeds."""
    return pack_policy("sampling", [("Karamouzas", 1.0, {"range": rng})],
                       sampling={"SamplingType": "regular", "SamplingBase": "zero",
                                 "SamplingBaseDirection": "preferred velocity", "SamplingAngle": 180,
                                 "SpeedSamples": speeds, "AngleSamples": angles,
                                 "SamplingRadius": "maximum speed", "IncludeBaseAsSample": "true"})


# -------------------------------------------------------------- workloads
def circle(n=50, radius=10.0, agent_radius=0.24, pref=1.3, vmax=1.6, max_neighbors=10, steps=300):
    """Circle crossing: n agents evenly spaced on a circle, goals at the antipodes;
    ORCA (time horizon 5 s, range 8 m), dt 0.1 s, infinite world."""
    a = _agents(n)
    th = 2.0 * np.pi * np.arange(n) / n
    a["pos"][:, 0] = (radius * np.cos(th)).astype(np.float32)
    a["pos"][:, 1] = (radius * np.sin(th)).astype(np.float32)
    a["goal"][:] = -a["pos"]
    a["radius"][:] = agent_radius
    a["pref_speed"][:] = pref
    a["max_speed"][:] = vmax
    a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
    return Scenario("circle%d" % n, 0, 0.0, 0.0, 0.1, max_neighbors, [orca_policy(5.0, 8.0)], a, steps)


def bidirectional(n=100_000, size=400.0, max_neighbors=16, steps=200, seed=2):
    """Dense bidirectional flow in a toric world: uniform positions, radius U[0.2,0.3],
    preferred speed clip(N(1.3,0.1),0.8,1.8), even IDs head +x and odd IDs -x with goals
    1000 m away; Karamouzas with 16x8 regular sampling, range 5 m, dt 0.1 s."""
    r = np.random.default_rng(seed)
    a = _agents(n)
    a["pos"][:] = r.uniform(-size / 2, size / 2, (n, 2)).astype(np.float32)
    a["radius"][:] = r.uniform(0.2, 0.3, n).astype(np.float32)
    a["pref_speed"][:] = np.clip(r.normal(1.3, 0.1, n), 0.8, 1.8).astype(np.float32)
    a["max_speed"][:] = 1.8
    sgn = np.where(np.arange(n) % 2 == 0, 1.0, -1.0).astype(np.float32)
    a["goal"][:, 0] = a["pos"][:, 0] + 1000.0 * sgn
    a["goal"][:, 1] = a["pos"][:, 1]
    a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
    return Scenario("bidirectional%d" % n, 1, size, size, 0.1, max_neighbors,
                    [karamouzas_sampling_policy(5.0, 16, 8)], a, steps)


def mixed(n=1_000_000, size=2000.0, max_neighbors=12, steps=100, seed=3):
    """Large mixed crowd in a toric world: uniform positions, goals 50-200 m away in a
    random direction; policy by ID mod 10: 0-3 PowerLaw + sampling (k 1.5, tau0 3 s,
    12x6 samples), 4-6 SocialForces + gradient (A=V0 2000, B=sigma 0.08), 7-9 ORCA
    (time horizon 3 s); range 4 m; dt 0.05 s."""
    r = np.random.default_rng(seed)
    a = _agents(n)
    a["pos"][:] = r.uniform(-size / 2, size / 2, (n, 2)).astype(np.float32)
    d = r.uniform(50.0, 200.0, n)
    phi = r.uniform(0.0, 2.0 * np.pi, n)
    a["goal"][:, 0] = (a["pos"][:, 0] + d * np.cos(phi)).astype(np.float32)
    a["goal"][:, 1] = (a["pos"][:, 1] + d * np.sin(phi)).astype(np.float32)
    m = np.arange(n) % 10
    a["policy"][:] = np.where(m < 4, 0, np.where(m < 7, 1, 2)).astype(np.int32)
    a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
    pols = [powerlaw_sampling_policy(1.5, 3.0, 4.0, 12, 6),
            socialforces_gradient_policy(2000.0, 0.08, 4.0),
            orca_policy(3.0, 4.0)]
    return Scenario("mixed%d" % n, 1, size, size, 0.05, max_neighbors, pols, a, steps)


def random_cloud(n, size, policies, world_type=1, max_neighbors=0, seed=0, dt=0.1, speed=1.0):
    """Small dense random crowd for parity tests: uniform positions, random goals,
    random initial velocities; agent i uses policy i mod len(policies)."""
    r = np.random.default_rng(seed)
    a = _agents(n)
    a["pos"][:] = r.uniform(-size / 2, size / 2, (n, 2)).astype(np.float32)
    a["goal"][:] = (a["pos"] + r.uniform(-20, 20, (n, 2))).astype(np.float32)
    a["vel"][:] = r.uniform(-speed, speed, (n, 2)).astype(np.float32)
    a["radius"][:] = r.uniform(0.2, 0.3, n).astype(np.float32)
    a["pref_speed"][:] = r.uniform(1.0, 1.6, n).astype(np.float32)
    a["policy"][:] = (np.arange(n) % len(policies)).astype(np.int32)
    a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
    return Scenario("cloud%d" % n, world_type, size, size, dt, max_neighbors, list(policies), a, 1)


@dataclass
class Batch:
    """Independent worlds stepped together (one policy per world here)."""
    name: str
    worlds: list                 # (world_type, width, height, dt) per world
    policies: list
    agents: dict                 # as Scenario.agents plus "world"
    max_neighbors: int
    steps: int
    sizes: np.ndarray

    @property
    def n(self):
        return len(self.agents["id"])


def sweep(n_worlds=512, n_min=20, n_max=5000, steps=500, seed=4, max_neighbors=0, dt=0.1):
    """Batched scenario sweep: agent counts log-uniform in [n_min, n_max]; layout by
    world index mod 3 (circle of circumference N metres with antipodal goals; two
    opposing lanes 4 m wide at 0.5 agents/m^2; uniform square at 0.5 agents/m^2 with
    random goals); per world a {PowerLaw + GoalReachingForce} sampling policy with
    k ~ U[1,3], tau0 ~ U[1,5] s, 12 angles x 6 speeds; infinite worlds; dt 0.1 s.
    Agent IDs restart at 0 in every world, as each world is its own simulator."""
    r = np.random.default_rng(seed)
    sizes = np.rint(np.exp(r.uniform(np.log(n_min), np.log(n_max), n_worlds))).astype(np.int64)
    parts, pols, worlds = [], [], []
    for w, nw in enumerate(sizes):
        nw = int(nw)
        a = _agents(nw)
        lay = w % 3
        if lay == 0:
            R = nw * 1.0 / (2.0 * np.pi)
            th = 2.0 * np.pi * np.arange(nw) / nw
            a["pos"][:, 0] = (R * np.cos(th)).astype(np.float32)
            a["pos"][:, 1] = (R * np.sin(th)).astype(np.float32)
            a["goal"][:] = -a["pos"]
        elif lay == 1:
            L = nw / (0.5 * 4.0)
            half = nw // 2
            x = r.uniform(0.0, L, nw)
            y = r.uniform(-2.0, 2.0, nw)
            left = np.arange(nw) < half
            a["pos"][:, 0] = np.where(left, -10.0 - x, 10.0 + x).astype(np.float32)
            a["pos"][:, 1] = y.astype(np.float32)
            a["goal"][:, 0] = np.where(left, a["pos"][:, 0] + L + 40.0, a["pos"][:, 0] - L - 40.0).astype(np.float32)
            a["goal"][:, 1] = a["pos"][:, 1]
        else:
            side = np.sqrt(nw / 0.5)
            a["pos"][:] = r.uniform(-side / 2, side / 2, (nw, 2)).astype(np.float32)
            a["goal"][:] = r.uniform(-side / 2, side / 2, (nw, 2)).astype(np.float32)
        a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
        a["policy"][:] = w
        a["world"] = np.full(nw, w, np.int32)
        k = float(r.uniform(1.0, 3.0))
        tau0 = float(r.uniform(1.0, 5.0))
        pols.append(powerlaw_sampling_policy(k=k, tau0=tau0, rng=5.0, angles=12, speeds=6))
        worlds.append((0, 0.0, 0.0, dt))
        parts.append(a)
    agents = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    return Batch("sweep%d" % n_worlds, worlds, pols, agents, max_neighbors, steps, sizes)


def powerlaw_gradient_policy(k=1.5, tau0=3.0, rng=5.0, relax=0.5):
    """{PowerLaw + GoalReachingForce} with the gradient method (examples/policies/PowerLaw.xml)."""
    return pack_policy("gradient", [("PowerLaw", 1.0, {"k": k, "tau0": tau0, "range": rng}),
                                    ("GoalReachingForce", 1.0, {})], relaxation_time=relax)


def calibration(n=10_000, size=100.0, V0=1500.0, sigma=0.12, seed=5, dt=0.1, steps=100, max_accel=5.0,
                max_neighbors=0, spread_deg=20.0, model="socialforces", k=1.5, tau0=3.0):
    """Adjoint calibration crowd: n agents in a size x size toric world at one agent per
    square metre, on a jittered lattice (spacing size / sqrt(n), jitter +-20% of it, so
    nobody starts overlapping), all walking in one general direction (+x within
    +-spread_deg) towards goals 100 m away; {SocialForcesAvoidance (A = V0, B = sigma,
    agentMass 80, range 5 m) + GoalReachingForce} with the gradient method, dt 0.1 s,
    the reference's default maximum acceleration. Initial velocities
    v0 = v_pref (1 + 0.1 N(0,1)). Random placement at this density starts half the
    agents in contact and the 5000 N/m contact forces make the rollout chaotic within
    a few steps (FP32 and FP64 trajectories part by metres within 40 steps); this
    lattice flow stays within about 1e-4 m for 20 steps. model="powerlaw": the same crowd
    with {PowerLaw (k, tau0, range 5 m) + GoalReachingForce} instead."""
    r = np.random.default_rng(seed)
    a = _agents(n)
    m = int(np.ceil(np.sqrt(n)))
    sp = size / m
    lat = (np.stack(np.meshgrid(np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 2)[:n] + 0.5) * sp - size / 2
    a["pos"][:] = (lat + r.uniform(-0.2 * sp, 0.2 * sp, lat.shape)).astype(np.float32)
    phi = np.deg2rad(r.uniform(-spread_deg, spread_deg, n))
    a["goal"][:, 0] = (a["pos"][:, 0] + 100.0 * np.cos(phi)).astype(np.float32)
    a["goal"][:, 1] = (a["pos"][:, 1] + 100.0 * np.sin(phi)).astype(np.float32)
    a["max_accel"][:] = max_accel
    d = a["goal"] - a["pos"]
    pref = d / np.linalg.norm(d, axis=1, keepdims=True) * a["pref_speed"][:, None]
    a["vel"][:] = (pref * (1.0 + 0.1 * r.standard_normal(n))[:, None]).astype(np.float32)
    a["viewdir"] = initial_viewdir(a["pos"], a["goal"])
    if model == "powerlaw":
        pol = powerlaw_gradient_policy(k=k, tau0=tau0, rng=5.0, relax=0.5)
        name = "calibration_powerlaw%d" % n
    else:
        pol = socialforces_gradient_policy(V0=V0, sigma=sigma, rng=5.0, mass=80.0, relax=0.5)
        name = "calibration%d" % n
    return Scenario(name, 1, size, size, dt, max_neighbors, [pol], a, steps)


# the two parameters the adjoint differentiates, per cost kind: parameter slots of
# (V0, sigma) of SocialForcesAvoidance and (k, tau0) of PowerLaw
PARAM_SLOTS = {2: (2, 3), 1: (1, 2)}


def with_params(packed, kind, p0, p1):
    """Copy of a packed policy with the two differentiated parameters of every cost of
    `kind` replaced: (V0, sigma) for SocialForcesAvoidance (2), (k, tau0) for PowerLaw (1)."""
    q = {k: (list(v) if isinstance(v, list) else v) for k, v in packed.items()}
    q["cost_params"] = [list(x) for x in packed["cost_params"]]
    s0, s1 = PARAM_SLOTS[kind]
    for c in range(q["n_cost"]):
        if q["cost_kind"][c] == kind:
            q["cost_params"][c][s0] = float(p0)
            q["cost_params"][c][s1] = float(p1)
    return q


def with_sf_params(packed, V0, sigma):
    """Copy of a packed policy with the SocialForcesAvoidance V0 and sigma replaced."""
    return with_params(packed, 2, V0, sigma)
