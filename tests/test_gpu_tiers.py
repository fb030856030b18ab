"""Long neighbour lists (no cap) have several launch plans: by default every agent's
neighbours are collected unsorted and ranked by its wave (list_capacity 1), agents with
more than 128 in range being redone by a 128-entry sorted insertion; or tiers, a launch
with 32- or 64-entry lists handing the agents whose list is longer to a launch with
longer lists. The plan must not show: every one gives the same lists and the same new
state, bit for bit, as one sorted-insertion launch at full capacity."""
import numpy as np
import pytest

import oracle_py as O
from cases import CASES
from gpu_util import REL_TOL, rel_err
from umans_amd import gpu
from umans_amd import scenarios as S
from umans_amd.registry import pack_policy

pytestmark = pytest.mark.gpu


def dense_world(case, world_type, list_capacity, size=50.0, steps=3):
    method, costs, samp, relax = CASES[case]
    # lists of every tier: ~48 neighbours on average (4 m range at 0.96 agents per m^2,
    # 5 m at 0.6), a few above 64 and a few at or below 32
    n = 2400 if case == "powerlaw_sampling" else 1500
    sc = S.random_cloud(n, size, [None], world_type=world_type, seed=31)
    w = gpu.World(world_type, size, size, sc.dt, [pack_policy(method, costs, sampling=samp, relaxation_time=relax)])
    w.upload(sc.agents)
    w.set_tuning(1.0, 0, list_capacity)
    w.record_neighbors(True)
    w.step(steps)
    out = w.download()
    nb = w.neighbors()
    st = w.stats()
    w.close()
    return out, nb, st


@pytest.mark.parametrize("case", ["powerlaw_sampling", "socialforces_gradient", "orca_global", "rvo_random"])
@pytest.mark.parametrize("world_type", [0, 1])
def test_list_capacity_tiers_do_not_change_results(case, world_type):
    ref, (rc, rid, rd2, rseg), rst = dense_world(case, world_type, 128)
    assert rst["max_list"] > 64, "the scenario must reach the last tier"
    for cap in (1, 32, 64):
        out, (cnt, ids, d2, seg), st = dense_world(case, world_type, cap)
        np.testing.assert_array_equal(cnt, rc)
        np.testing.assert_array_equal(ids, rid)
        np.testing.assert_array_equal(d2, rd2)
        for k in ("pos", "vel", "viewdir", "rng"):
            np.testing.assert_array_equal(out[k], ref[k], err_msg=f"{case} capacity {cap} {k}")
        assert st["overflow_agents"] == rst["overflow_agents"]
        assert st["max_list"] == rst["max_list"]


@pytest.mark.parametrize("world_type", [0, 1])
def test_ranked_lists_beyond_their_capacity(world_type):
    """With a cap of 128, agents with more than 128 neighbours in range leave the ranked
    lists for the 128-entry sorted insertion, which keeps the nearest 128: the same lists
    and state as a single sorted-insertion launch, and the overflow counted once. (With no
    cap such lists are kept whole: tests/test_gpu_dense.py.)"""
    def run(list_capacity):
        method, costs, samp, relax = CASES["powerlaw_sampling"]
        sc = S.random_cloud(2200, 25.0, [None], world_type=world_type, seed=32)
        w = gpu.World(world_type, 25.0, 25.0, sc.dt, [pack_policy(method, costs, sampling=samp,
                                                                   relaxation_time=relax)], max_neighbors=128)
        w.upload(sc.agents)
        w.set_tuning(1.0, 0, list_capacity)
        w.record_neighbors(True)
        w.step(2)
        out, nb, st = w.download(), w.neighbors(), w.stats()
        w.close()
        return out, nb, st
    ref, (rc, rid, rd2, rseg), rst = run(128)
    assert rst["overflow_agents"] > 0, "some lists must exceed 128"
    out, (cnt, ids, d2, seg), st = run(1)
    np.testing.assert_array_equal(cnt, rc)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(d2, rd2)
    np.testing.assert_array_equal(seg, rseg)
    for k in ("pos", "vel", "viewdir", "rng"):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    assert st["overflow_agents"] == rst["overflow_agents"]
    assert st["max_list"] == rst["max_list"]


def lattice(side, spacing, jitter, seed=5):
    """Agents on a square lattice; with jitter, each coordinate moved by 0 or +-2^-13 m,
    so that many distances in a list round to the same float but differ as doubles."""
    g = (np.arange(side, dtype=np.float32) - side / 2 + 0.5) * np.float32(spacing)
    xx, yy = np.meshgrid(g, g)
    pos = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float32)
    if jitter:
        rng = np.random.default_rng(seed)
        pos += rng.choice(np.float32([0.0, 2.0 ** -13, -(2.0 ** -13)]), pos.shape).astype(np.float32)
    return pos


def float_only_ties(pos, rng_m):
    """Number of agents whose in-range list holds two distances equal as floats and
    unequal as doubles (infinite world)."""
    p = pos.astype(np.float64)
    n = 0
    for i in range(len(p)):
        d2 = ((p - p[i]) ** 2).sum(1)
        d2 = d2[(d2 < np.float64(np.float32(rng_m * rng_m))) & (np.arange(len(p)) != i)]
        f = d2.astype(np.float32)
        o = np.argsort(d2, kind="stable")
        d2, f = d2[o], f[o]
        n += bool(np.any((f[1:] == f[:-1]) & (d2[1:] != d2[:-1])))
    return n


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("spacing", [0.75, 1.5, 2.5])
@pytest.mark.parametrize("jitter", [False, True])
def test_ranked_lists_on_a_lattice(world_type, spacing, jitter):
    """Agents on a square lattice (range 4 m; spacings giving lists of ~89, ~22 and ~8
    entries, so rows of every size the wave sorts): every list holds many equal distances, ordered by
    agent ID; with jitter, also distances equal as floats and unequal as doubles, ordered
    by the doubles. The ranked lists (sorted by packed keys, then again by exact doubles
    and IDs where floats tie) equal the sorted insertion's, bit for bit."""
    side = 36
    pos = lattice(side, spacing, jitter)
    if jitter:
        assert float_only_ties(pos[:100], 4.0) > 0, "the jitter must make float-only ties"

    def run(list_capacity):
        method, costs, samp, relax = CASES["powerlaw_sampling"]
        sc = S.random_cloud(side * side, float(side * spacing), [None], world_type=world_type, seed=33)
        a = dict(sc.agents)
        a["pos"] = pos
        rng = np.random.default_rng(5)
        a["id"] = rng.permutation(side * side).astype(np.int32)   # IDs unrelated to positions
        w = gpu.World(world_type, float(side * spacing), float(side * spacing), sc.dt,
                      [pack_policy(method, costs, sampling=samp, relaxation_time=relax)])
        w.upload(a)
        w.set_tuning(1.0, 0, list_capacity)
        w.record_neighbors(True)
        w.step(1)
        out, nb, st = w.download(), w.neighbors(), w.stats()
        w.close()
        return out, nb, st
    ref, (rc, rid, rd2, rseg), rst = run(128)
    out, (cnt, ids, d2, seg), st = run(1)
    np.testing.assert_array_equal(cnt, rc)
    np.testing.assert_array_equal(ids, rid)
    np.testing.assert_array_equal(d2, rd2)
    np.testing.assert_array_equal(seg, rseg)
    for k in ("pos", "vel", "viewdir", "rng"):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


@pytest.mark.parametrize("world_type", [0, 1])
@pytest.mark.parametrize("spacing", [0.75, 1.5])
@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("cap", [12, 16])
def test_capped_short_lists_on_a_lattice(world_type, spacing, jitter, cap):
    """Capped lists (12 or 16, the per-lane LDS buffers) on the square lattice: lists of ~89
    or ~22 agents in range, so every buffer fills and drops its canonical last entry many
    times, and rings of 4 or 8 equal distances straddle the cut (with jitter, distances
    equal as floats and unequal as doubles too). Lists bit-exact against the FP32 oracle
    (the first `cap` of the canonical order: exact d2, then agent ID), velocities to the
    parity bar."""
    side = 36
    pos = lattice(side, spacing, jitter, seed=7)
    pol = S.powerlaw_sampling_policy()
    sc = S.random_cloud(side * side, float(side * spacing), [pol], world_type=world_type, seed=41)
    a = dict(sc.agents)
    a["pos"] = pos
    a["id"] = np.random.default_rng(9).permutation(side * side).astype(np.int32)
    w = gpu.World(world_type, float(side * spacing), float(side * spacing), sc.dt, [pol], max_neighbors=cap)
    w.upload(a)
    w.record_neighbors(True)
    w.step(1)
    out, (cnt, ids, d2, seg), st = w.download(), w.neighbors(), w.stats()
    w.close()
    wd = O.make_world(world_type, float(side * spacing), float(side * spacing), sc.dt, cap)
    o32 = O.step(wd, [pol], a, precision=0, list_stride=128)
    o64 = O.step(wd, [pol], a, precision=1)
    np.testing.assert_array_equal(cnt, o32["nb_count"])
    m = np.arange(ids.shape[1])[None] < cnt[:, None]
    np.testing.assert_array_equal(np.where(m, ids, -1), np.where(m, o32["nb_id"][:, :ids.shape[1]], -1))
    np.testing.assert_array_equal(np.where(m, d2, 0.0), np.where(m, o32["nb_d2"][:, :ids.shape[1]], 0.0))
    np.testing.assert_array_equal(np.where(m, seg, 0), np.where(m, o32["nb_seg"][:, :ids.shape[1]], 0))
    assert st["overflow_agents"] == int((o32["nb_full"] > cap).sum()) > 0
    e64 = rel_err(out["vel"], o64["vel64"])
    e32 = rel_err(out["vel"], o32["vel"])
    ok = (e64 <= REL_TOL) | (e32 <= 1e-6)
    assert ok.all(), f"{(~ok).sum()} agents off: e64 {np.sort(e64)[-4:]}"
