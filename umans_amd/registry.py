"""Cost-function registry and policy packing (host side).

Mirrors CostFunctionFactory (src/Engine/core/costFunctionFactory.cpp:41-61) and the
parameter readers of each cost function's parseParameters: the same registry names,
XML attribute names and defaults. A policy is packed into the flat numeric layout of
`umans_policy` (include/umans_gpu.h). The C++ side of the library holds the same
table (umans_amd/csrc/registry.cpp); tests check that both agree.
"""
import math

METHODS = {"gradient": 0, "sampling": 1, "global": 2}             # policy.cpp:151-162
SAMPLING_TYPE = {"regular": 0, "random": 1}                        # policy.cpp:187-197
SAMPLING_BASE = {"zero": 0, "current velocity": 1}                 # policy.cpp:164-173
SAMPLING_BASE_DIR = {"unit": 0, "current velocity": 1, "preferred velocity": 2}   # policy.cpp:175-185
SAMPLING_RADIUS = {"preferred speed": 0, "maximum speed": 1, "maximum acceleration": 2}  # policy.cpp:199-210

# name -> (kind, default range, [(xml attribute, slot, default)] in the reference's read order)
# Defaults: CostFunction::range_ 5 (costFunction.h:132); GoalReachingForce range 0
# (GoalReachingForce.h:38); agentMass 1 (ForceBasedFunction.h:39); the rest from each header.
COSTS = {
    "GoalReachingForce": (0, 0.0, [("agentMass", 0, 1.0)]),
    "PowerLaw": (1, 5.0, [("agentMass", 0, 1.0), ("k", 1, 1.5), ("tau0", 2, 3.0)]),
    # SocialForcesAvoidance.cpp:142-150 reads "U0" into V0 after "V0" (slot 2 twice).
    "SocialForcesAvoidance": (2, 5.0, [("agentMass", 0, 1.0), ("dt", 1, 1.0), ("V0", 2, 2.1),
                                       ("sigma", 3, 0.3), ("U0", 2, None), ("R", 4, 0.2)]),
    "ORCA": (3, 5.0, [("timeHorizon", 0, 5.0)]),
    "Karamouzas": (4, 5.0, [("alpha", 0, 5.0), ("beta", 1, 0.5), ("gamma", 2, 1.0), ("delta", 3, 1.0),
                            ("t_max", 4, 8.0)]),
    "RVO": (5, 5.0, [("w", 0, 1.0)]),
    "Moussaid": (6, 5.0, [("d_max", 0, 8.0)]),
    "TtcaDca": (7, 5.0, [("sigmaAngle_goal", 0, 2.0), ("sigmaSpeed_goal", 1, 2.0), ("sigmaTtca", 2, 1.0),
                         ("sigmaDca", 3, 0.3)]),
    "Paris": (8, 5.0, [("w_a", 0, 0.5), ("w_b", 1, 0.0), ("t_max", 2, 8.0)]),
    "VanToll": (9, 5.0, [("max_distance", 0, 8.0)]),
}
KIND_NAMES = {v[0]: k for k, v in COSTS.items()}


def cost_defaults(name):
    kind, rng, table = COSTS[name]
    params = [0.0] * 8
    for _, slot, dflt in table:
        if dflt is not None:
            params[slot] = float(dflt)
    return kind, rng, params


def pack_policy(method="gradient", costs=(), sampling=None, relaxation_time=0.0,
                contact_force_scale=5000.0, gradient_iterations=1, gradient_step=1.0):
    """Build a packed policy (dict with the umans_policy fields).

    costs: sequence of (name, coeff, {param: value}) — `range` may appear among the params.
    sampling: dict of SamplingParameters using the XML attribute names
      (SamplingType, SamplingBase, SamplingBaseDirection, SamplingRadius, SamplingAngle,
       SpeedSamples, AngleSamples, RandomSamples, IncludeBaseAsSample); defaults from
      policy.h:63-71. Defaults of RelaxationTime 0 and ContactForceScale 5000: policy.h:121-125.
    gradient_iterations, gradient_step: the gradient method's descent extension (not in
      the reference; 1 keeps its single step): v <- v - gradient_step * dt * sum coeff grad C(v),
      gradient_iterations times, then acceleration (v_N - v) / dt.
    """
    s = dict(sampling or {})
    p = {
        "method": METHODS[method] if isinstance(method, str) else int(method),
        "s_type": SAMPLING_TYPE[s.get("SamplingType", "regular")],
        "s_base": SAMPLING_BASE[s.get("SamplingBase", "zero")],
        "s_base_dir": SAMPLING_BASE_DIR[s.get("SamplingBaseDirection", "current velocity")],
        "s_radius": SAMPLING_RADIUS[s.get("SamplingRadius", "preferred speed")],
        "s_angle": float(s.get("SamplingAngle", 180.0)),
        "s_speed_samples": int(s.get("SpeedSamples", 4)),
        "s_angle_samples": int(s.get("AngleSamples", 11)),
        "s_random_samples": int(s.get("RandomSamples", 100)),
        "s_include_base": int(str(s.get("IncludeBaseAsSample", "false")).lower() in ("1", "true")),
        "relaxation_time": float(relaxation_time),
        "contact_force_scale": float(contact_force_scale),
        "gradient_iterations": int(gradient_iterations), "gradient_step": float(gradient_step),
        "n_cost": 0, "cost_kind": [], "cost_coeff": [], "cost_range": [], "cost_params": [],
    }
    if len(costs) > 4:
        raise ValueError("at most 4 cost functions per policy")
    for name, coeff, params in costs:
        if name not in COSTS:
            raise KeyError(f"unknown cost function {name!r}")
        kind, rng, vals = cost_defaults(name)
        table = COSTS[name][2]
        params = dict(params or {})
        if "range" in params:
            rng = float(params.pop("range"))
        for attr, slot, _ in table:            # reference read order (U0 after V0)
            if attr in params:
                vals[slot] = float(params[attr])
        unknown = set(params) - {a for a, _, _ in table}
        if unknown:
            raise KeyError(f"unknown parameters for {name}: {sorted(unknown)}")
        p["cost_kind"].append(kind)
        p["cost_coeff"].append(float(coeff))
        p["cost_range"].append(rng)
        p["cost_params"].append(vals)
        p["n_cost"] += 1
    for _ in range(4 - p["n_cost"]):
        p["cost_kind"].append(0); p["cost_coeff"].append(0.0); p["cost_range"].append(0.0)
        p["cost_params"].append([0.0] * 8)
    return p


def interaction_range(packed):
    """Policy::getInteractionRange (policy.cpp:34-41)."""
    r = 0.0
    for c in range(packed["n_cost"]):
        r = max(r, packed["cost_range"][c])
    return r


def to_reference_xml_policy(method, costs, sampling=None, relaxation_time=None, contact_force_scale=None):
    """The same policy as the dict the reference's XML loader reads (crowdSimulator.cpp:327-422)."""
    d = {"method": method, "costs": []}
    if sampling:
        d["sampling"] = dict(sampling)
    if relaxation_time is not None:
        d["relaxation_time"] = relaxation_time
    if contact_force_scale is not None:
        d["contact_force_scale"] = contact_force_scale
    for name, coeff, params in costs:
        attrs = {"coeff": coeff}
        attrs.update(params or {})
        d["costs"].append((name, attrs))
    return d


def to_reference_xml_from_packed(packed):
    """A packed policy as the reference XML loader's policy (crowdSimulator.cpp:327-422):
    the optimisation method and SamplingParameters attributes, RelaxationTime,
    ContactForceScale and each cost function with its coefficient, range and parameters
    (registry attribute names; "U0" is not written: it only aliases V0)."""
    inv = lambda table, v: [k for k, x in table.items() if x == v][0]
    method = inv(METHODS, packed["method"])
    sampling = None
    if method != "gradient":
        sampling = {"SamplingType": inv(SAMPLING_TYPE, packed["s_type"]),
                    "SamplingBase": inv(SAMPLING_BASE, packed["s_base"]),
                    "SamplingBaseDirection": inv(SAMPLING_BASE_DIR, packed["s_base_dir"]),
                    "SamplingRadius": inv(SAMPLING_RADIUS, packed["s_radius"]),
                    "SamplingAngle": float(packed["s_angle"]), "SpeedSamples": int(packed["s_speed_samples"]),
                    "AngleSamples": int(packed["s_angle_samples"]), "RandomSamples": int(packed["s_random_samples"]),
                    "IncludeBaseAsSample": "true" if packed["s_include_base"] else "false"}
    costs = []
    for c in range(packed["n_cost"]):
        name = KIND_NAMES[packed["cost_kind"][c]]
        params = {"range": float(packed["cost_range"][c])} if name != "GoalReachingForce" else {}
        for attr, slot, _ in COSTS[name][2]:
            if attr != "U0":
                params[attr] = float(packed["cost_params"][c][slot])
        costs.append((name, float(packed["cost_coeff"][c]), params))
    return to_reference_xml_policy(method, costs, sampling=sampling,
                                   relaxation_time=float(packed["relaxation_time"]),
                                   contact_force_scale=float(packed["contact_force_scale"]))
