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
