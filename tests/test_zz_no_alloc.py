"""umans_step_noalloc, the entry point that allocates no device memory, against umans_step
on a context: the same world, policies and agents give bit-identical states. Two calls
in a row, the second with fewer agents (the fixed scratch array keeps the first call's
contents), each through the C ABI on device arrays the test allocates."""
import ctypes as C

import numpy as np
import pytest
import torch

from umans_amd import gpu, registry, scenarios as S

pytestmark = pytest.mark.gpu

OUT_FIELDS = ("pos", "vel", "viewdir", "rng", "pref_vel")


def policy_mix():
    """One policy per step-kernel class: force + sampling, force + gradient, ORCA,
    generic (random sampling draws from the per-agent RNG) and cooperative sampling."""
    return [S.powerlaw_sampling_policy(1.5, 3.0, 4.0, 12, 6),
            S.socialforces_gradient_policy(2000.0, 0.08, 4.0),
            S.orca_policy(3.0, 4.0),
            registry.pack_policy("sampling", [("RVO", 1.0, {"range": 4.0})],
                                 sampling={"SamplingType": "random", "RandomSamples": 40,
                                           "SamplingRadius": "maximum speed", "SamplingAngle": 360}),
            S.karamouzas_sampling_policy(4.0, 8, 4)]


def context_run(sc, nsteps):
    W = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, max_neighbors=sc.max_neighbors)
    try:
        W.upload(sc.agents)
        W.step(nsteps)
        return W.download()
    finally:
        W.close()


def noalloc_run(sc, nsteps, explicit_rng):
    dev = torch.device("cuda")
    a = sc.agents
    n = sc.n
    t = {
        "id": torch.tensor(a["id"], dtype=torch.int32, device=dev),
        "pos": torch.tensor(a["pos"], dtype=torch.float32, device=dev),
        "vel": torch.tensor(a["vel"], dtype=torch.float32, device=dev),
        "goal": torch.tensor(a["goal"], dtype=torch.float32, device=dev),
        "viewdir": torch.tensor(a["viewdir"], dtype=torch.float32, device=dev),
        "radius": torch.tensor(a["radius"], dtype=torch.float32, device=dev),
        "pref_speed": torch.tensor(a["pref_speed"], dtype=torch.float32, device=dev),
        "max_speed": torch.tensor(a["max_speed"], dtype=torch.float32, device=dev),
        "max_accel": torch.tensor(a["max_accel"], dtype=torch.float32, device=dev),
        "policy": torch.tensor(a["policy"], dtype=torch.int32, device=dev),
        "pref_vel": torch.zeros((n, 2), dtype=torch.float32, device=dev),
    }
    if explicit_rng:
        # std::default_random_engine(id) seeds (agent.cpp:38), as umans_upload seeds them
        seeds = (a["id"].astype(np.uint32) % np.uint32(2147483647)).astype(np.uint32)
        seeds[seeds == 0] = 1
        t["rng"] = torch.tensor(seeds.view(np.int32), dtype=torch.int32, device=dev)
    else:
        t["rng"] = None
    ptrs = {k: v.data_ptr() for k, v in t.items() if v is not None}
    stream = torch.cuda.current_stream()
    gpu.step_noalloc(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors, ptrs, n, nsteps,
                     stream=stream.cuda_stream)
    stream.synchronize()
    out = {k: t[k].cpu().numpy() for k in ("pos", "vel", "viewdir", "pref_vel")}
    if t["rng"] is not None:
        out["rng"] = t["rng"].cpu().numpy().view(np.uint32)
    return out


def assert_same(got, ref, what):
    for k in OUT_FIELDS:
        if k not in got:
            continue
        assert np.array_equal(got[k], ref[k]), f"{what}: {k} differs at " \
            f"{np.argwhere(np.any(np.atleast_2d(got[k].T).T != np.atleast_2d(ref[k].T).T, axis=-1))[:5].ravel()}"


def test_noalloc_matches_context_steps_twice():
    # first call: 3000 agents in a toric world, lists capped at 16, 3 steps, RNG seeded
    # on the device from the IDs
    sc1 = S.random_cloud(3000, 70.0, policy_mix(), world_type=1, max_neighbors=16, seed=21)
    ref1 = context_run(sc1, 3)
    got1 = noalloc_run(sc1, 3, explicit_rng=False)
    assert_same(got1, ref1, "first call")
    # second call, smaller: 1200 agents in an infinite world, lists capped at 64 (the
    # 64-entry sorted insertion), 2 steps, explicit RNG state
    sc2 = S.random_cloud(1200, 30.0, policy_mix(), world_type=0, max_neighbors=64, seed=22)
    ref2 = context_run(sc2, 2)
    got2 = noalloc_run(sc2, 2, explicit_rng=True)
    assert_same(got2, ref2, "second call")
    assert not np.array_equal(got1["pos"][:sc2.n], got2["pos"])


def test_noalloc_rejects_sizes_beyond_the_maximum():
    L = gpu.lib()
    desc = gpu.WorldDesc(1, 100.0, 100.0, 0.1, 12, 0.0, 0, 0)
    pols = (gpu.PolicyC * 1)(gpu.policy_struct(S.orca_policy(3.0, 4.0)))
    a = gpu.AgentsC()
    err = C.create_string_buffer(128)
    rc = L.umans_step_noalloc(C.byref(desc), pols, 1, C.byref(a), 16385, 1, None, err, len(err))
    assert rc == -22 and b"UMANS_NOALLOC_MAX_AGENTS" in err.value
    desc.max_neighbors = 0                         # uncapped lists are not supported here
    rc = L.umans_step_noalloc(C.byref(desc), pols, 1, C.byref(a), 10, 1, None, err, len(err))
    assert rc == -22
