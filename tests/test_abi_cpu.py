"""CPU checks of the C ABI library: it loads, exports every function include/umans_gpu.h
declares, and its registry agrees with the reference's names, parameters and defaults
(the Python table in umans_amd.registry restates the same headers)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from umans_amd import gpu, registry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "umans_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(umans_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    L = gpu.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(gpu.EXPORTS)


def test_registry_names_match():
    L = gpu.lib()
    for name, (kind, rng, table) in registry.COSTS.items():
        assert L.umans_cost_id(name.encode()) == kind
        assert L.umans_cost_name(kind).decode() == name
    assert L.umans_cost_id(b"SocialForces") == -1        # stale name in test/Generator.py:130
    assert L.umans_cost_id(b"GoalReaching") == -1        # stale name in test/ConfigGenerator.py:8


def test_registry_params_and_defaults_match():
    L = gpu.lib()
    for name, (kind, rng, table) in registry.COSTS.items():
        r = C.c_float()
        p = (C.c_float * 8)()
        assert L.umans_cost_defaults(kind, C.byref(r), p) == 0
        _, prng, pvals = registry.cost_defaults(name)
        assert r.value == pytest.approx(prng)
        np.testing.assert_allclose(np.array(p[:]), np.array(pvals, np.float32))
        for attr, slot, _ in table:
            assert L.umans_cost_param_index(kind, attr.encode()) == slot
    # SocialForcesAvoidance reads "U0" into V0 (SocialForcesAvoidance.cpp:148)
    assert L.umans_cost_param_index(2, b"U0") == L.umans_cost_param_index(2, b"V0")
    assert L.umans_cost_param_name(2, 2).decode() == "V0"


def test_policy_init_defaults():
    L = gpu.lib()
    q = gpu.PolicyC()
    assert L.umans_policy_init(C.byref(q), 1) == 0
    ref = registry.pack_policy("sampling", [])
    for k in ("s_type", "s_base", "s_base_dir", "s_radius", "s_speed_samples", "s_angle_samples",
              "s_random_samples", "s_include_base"):
        assert getattr(q, k) == ref[k], k
    assert q.s_angle == ref["s_angle"] and q.contact_force_scale == 5000.0 and q.relaxation_time == 0.0


def test_pack_policy_u0_overrides_v0():
    p = registry.pack_policy("gradient", [("SocialForcesAvoidance", 1.0, {"V0": 3.0, "U0": 7.0})])
    assert p["cost_params"][0][2] == 7.0
    with pytest.raises(KeyError):
        registry.pack_policy("gradient", [("NoSuchCost", 1.0, {})])
