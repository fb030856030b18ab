"""Policy cases shared by the golden-fixture generator and the parity tests.

Each case is (method, [(registry name, coeff, params)], sampling attributes,
RelaxationTime), in the reference's XML vocabulary (crowdSimulator.cpp:327-422).
"""
SAMP = {"SamplingType": "regular", "SamplingBase": "zero", "SamplingBaseDirection": "preferred velocity",
        "SamplingAngle": 180, "SpeedSamples": 5, "AngleSamples": 9, "SamplingRadius": "maximum speed",
        "IncludeBaseAsSample": "true"}
SAMP360 = dict(SAMP, SamplingAngle=360, IncludeBaseAsSample="false")
RANDOM = {"SamplingType": "random", "SamplingBase": "current velocity", "SamplingBaseDirection": "unit",
          "SamplingAngle": 360, "RandomSamples": 50, "SamplingRadius": "maximum acceleration"}

CASES = {
    "orca_global": ("global", [("ORCA", 1.0, {"timeHorizon": 5.0, "range": 5})], None, 0.0),
    "orca_sampling": ("sampling", [("ORCA", 1.0, {"range": 5})], SAMP, 0.0),
    "orca_gradient": ("gradient", [("ORCA", 1.0, {"range": 5})], None, 0.0),
    "powerlaw_sampling": ("sampling", [("PowerLaw", 1.0, {"range": 4}), ("GoalReachingForce", 1.0, {})], SAMP360, 0.5),
    "powerlaw_gradient": ("gradient", [("PowerLaw", 1.0, {"range": 5}), ("GoalReachingForce", 1.0, {})], None, 0.5),
    "powerlaw_global": ("global", [("PowerLaw", 1.0, {"range": 4, "k": 2.0, "tau0": 2.5}), ("GoalReachingForce", 1.0, {})], None, 0.5),
    "socialforces_gradient": ("gradient", [("SocialForcesAvoidance", 1.0, {"range": 5, "V0": 2000, "sigma": 0.08, "agentMass": 80}), ("GoalReachingForce", 1.0, {})], None, 0.5),
    "socialforces_default": ("gradient", [("SocialForcesAvoidance", 1.0, {"range": 5}), ("GoalReachingForce", 1.0, {})], None, 0.5),
    "socialforces_u0": ("gradient", [("SocialForcesAvoidance", 1.0, {"range": 5, "V0": 3.0, "U0": 7.0}), ("GoalReachingForce", 1.0, {})], None, 0.5),
    "karamouzas_sampling": ("sampling", [("Karamouzas", 1.0, {"range": 5})], SAMP, 0.0),
    "karamouzas_global": ("global", [("Karamouzas", 1.0, {"range": 5})], None, 0.0),
    "rvo_random": ("sampling", [("RVO", 1.0, {"range": 5})], RANDOM, 0.0),
    "moussaid_sampling": ("sampling", [("Moussaid", 1.0, {"range": 5})], dict(SAMP, IncludeBaseAsSample="false"), 0.5),
    "ttcadca_gradient": ("gradient", [("TtcaDca", 20.0, {"range": 5})], None, 0.0),
    "ttcadca_sampling": ("sampling", [("TtcaDca", 1.0, {"range": 5})], SAMP, 0.0),
    "paris_sampling": ("sampling", [("Paris", 1.0, {"range": 5})], SAMP, 0.0),
    "vantoll_sampling": ("sampling", [("VanToll", 1.0, {"range": 5})], SAMP, 0.0),
    "vantoll_gradient": ("gradient", [("VanToll", 1.0, {"range": 5})], None, 0.0),
    "multi_cost_global": ("global", [("ORCA", 1.0, {"range": 5}), ("GoalReachingForce", 0.5, {})], None, 0.5),
    "goal_only_gradient": ("gradient", [("GoalReachingForce", 1.0, {})], None, 0.5),
}

# Cases re-run with obstacles (fixtures obst_<case>_w<world>): every model that reads
# obstacle edges (time-to-collision users, SocialForces' obstacle force) and two that
# see obstacles only through the contact forces.
OBSTACLE_CASES = ["karamouzas_sampling", "karamouzas_global", "rvo_random", "moussaid_sampling",
                  "vantoll_sampling", "vantoll_gradient", "socialforces_gradient", "socialforces_r",
                  "powerlaw_gradient", "orca_global", "paris_sampling", "ttcadca_sampling"]
CASES["socialforces_r"] = ("gradient", [("SocialForcesAvoidance", 1.0, {"range": 5, "R": 0.5}),
                                        ("GoalReachingForce", 1.0, {})], None, 0.5)

# Polygons (vertex lists) inside the 12 m test world: a square, a triangle, a wall given
# as a two-point polygon (two coincident edges), a one-point polygon (an edge of zero
# length) and a wall next to the x seam of the torus.
OBSTACLES = [[(-1.0, -1.0), (1.0, -1.0), (1.0, 1.0), (-1.0, 1.0)],
             [(3.0, 3.0), (5.0, 3.0), (4.0, 5.0)],
             [(-5.0, -3.0), (-2.0, -4.5)],
             [(-4.0, 3.0)],
             [(5.6, -2.0), (5.9, 2.0)]]
