"""Batched worlds, hipGraph replay and the autotuner, through the C ABI.

* a batch of worlds gives exactly the results of each world run alone (the same
  kernels see the same data), and each world matches the oracle;
* a step captured in a hipGraph and replayed gives exactly the eager results, replay
  after replay, with eager work and synchronisations in between;
* the autotuner leaves the state untouched and its choices change no result.
"""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import REL_TOL, rel_err
from umans_amd import gpu, scenarios as S

pytestmark = pytest.mark.gpu


def small_sweep():
    return S.sweep(n_worlds=12, n_min=20, n_max=400, seed=40)


def world_slice(b, w):
    sel = b.agents["world"] == w
    a = {k: v[sel] for k, v in b.agents.items() if k != "world"}
    a["policy"] = np.zeros(sel.sum(), np.int32)
    return a


def test_batch_equals_separate_worlds():
    b = small_sweep()
    W = gpu.World(0, 0, 0, 0.1, b.policies, max_neighbors=b.max_neighbors, batch=b.worlds)
    W.upload(b.agents)
    W.step(5)
    out = W.download()
    W.close()
    for w, (t, wd, ht, dt) in enumerate(b.worlds):
        a = world_slice(b, w)
        single = gpu.World(t, wd, ht, dt, [b.policies[w]], max_neighbors=b.max_neighbors)
        single.upload(a)
        single.step(5)
        o = single.download()
        single.close()
        sel = b.agents["world"] == w
        np.testing.assert_array_equal(out["pos"][sel], o["pos"], err_msg=f"world {w}")
        np.testing.assert_array_equal(out["vel"][sel], o["vel"], err_msg=f"world {w}")
        np.testing.assert_array_equal(out["world"][sel], w)


def test_batch_one_step_matches_oracle():
    b = small_sweep()
    W = gpu.World(0, 0, 0, 0.1, b.policies, max_neighbors=b.max_neighbors, batch=b.worlds)
    W.upload(b.agents)
    W.step(1)
    out = W.download()
    W.close()
    for w, (t, wd, ht, dt) in enumerate(b.worlds):
        a = world_slice(b, w)
        wdesc = O.make_world(t, wd, ht, dt, b.max_neighbors)
        o32 = O.step(wdesc, [b.policies[w]], a, precision=0)
        o64 = O.step(wdesc, [b.policies[w]], a, precision=1)
        sel = b.agents["world"] == w
        e64 = rel_err(out["vel"][sel], o64["vel64"])
        e32 = rel_err(out["vel"][sel], o32["vel"])
        assert ((e64 <= REL_TOL) | (e32 <= 1e-6)).all(), f"world {w}"


@pytest.mark.parametrize("kind", ["mixed", "mixed_eager_first", "sweep"])
def test_graph_replay_matches_eager(kind):
    """Replays of the captured step equal eager steps. With "mixed_eager_first" the graph
    is captured after an eager step, so it holds no clear of the cell counters and every
    replay relies on the scan leaving them zero."""
    if kind.startswith("mixed"):
        sc = S.mixed(n=20_000, size=283.0)
        mk = lambda: gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
        agents = sc.agents
    else:
        b = small_sweep()
        mk = lambda: gpu.World(0, 0, 0, 0.1, b.policies, max_neighbors=b.max_neighbors, batch=b.worlds)
        agents = b.agents
    eager = mk()
    eager.upload(agents)
    graph = mk()
    graph.upload(agents)
    if kind == "mixed_eager_first":
        eager.step(1)
        graph.step(1)
    graph.set_graph(True)
    for r in range(4):
        eager.step(2)
        graph.step(2)
        e = eager.download()      # eager work and synchronisations between replays
        g = graph.download()
        np.testing.assert_array_equal(g["pos"], e["pos"], err_msg=f"replay {r}")
        np.testing.assert_array_equal(g["vel"], e["vel"], err_msg=f"replay {r}")
        np.testing.assert_array_equal(g["viewdir"], e["viewdir"], err_msg=f"replay {r}")
    eager.close()
    graph.close()


def test_autotune_keeps_state_and_results():
    sc = S.mixed(n=50_000, size=447.0)
    ref = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    ref.upload(sc.agents)
    ref.step(3)
    r = ref.download()
    ref.close()
    W = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    W.upload(sc.agents)
    before = W.download()
    t = W.autotune(trial_steps=2)
    assert t["cell_factor"] > 0 and t["block_size"] in (64, 128, 256)
    after = W.download()
    np.testing.assert_array_equal(before["pos"], after["pos"])
    np.testing.assert_array_equal(before["vel"], after["vel"])
    W.step(3)
    o = W.download()
    np.testing.assert_array_equal(o["pos"], r["pos"])
    np.testing.assert_array_equal(o["vel"], r["vel"])
    for f, bsz in ((0.5, 64), (1.5, 256), (0.75, 128)):
        W.set_tuning(f, bsz)
        W.upload(sc.agents)
        W.step(3)
        o = W.download()
        np.testing.assert_array_equal(o["vel"], r["vel"], err_msg=f"tuning {f} {bsz}")
    W.close()


@pytest.mark.parametrize("kind", ["mixed", "toric_batch", "dynamic", "infinite_batch"])
def test_steps_in_one_call_equal_single_steps(kind):
    """Within one step call every step also counts the agents for the next one (fixed
    grids), and the next step skips its own count; with infinite worlds it fills the
    bounding boxes the next step's grids come from (one step old, where a single-step
    call measures them afresh). A call of K steps gives exactly the results of K calls of
    one step, with scheduled insertions and remove_at_goal removals in between for
    "dynamic"."""
    import os
    from goldens import GOLDEN
    if kind == "dynamic":
        path = os.path.join(GOLDEN, "examples", "Dynamic-Mixed.xml")
        mk = lambda: gpu.World.from_scenario_file(path)
        agents = None
        steps = 40
    elif kind == "infinite_batch":
        b = small_sweep()
        mk = lambda: gpu.World(0, 0, 0, 0.1, b.policies, max_neighbors=b.max_neighbors, batch=b.worlds)
        agents = b.agents
        steps = 12
    elif kind == "mixed":
        sc = S.mixed(n=20_000, size=283.0)
        mk = lambda: gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
        agents = sc.agents
        steps = 6
    else:
        sc = S.mixed(n=6_000, size=160.0)
        half = sc.n // 2
        worlds = [(1, sc.width, sc.height, sc.dt)] * 2
        a = {k: v.copy() for k, v in sc.agents.items()}
        a["world"] = (np.arange(sc.n) >= half).astype(np.int32)
        mk = lambda: gpu.World(0, 0, 0, sc.dt, sc.policies, max_neighbors=sc.max_neighbors, batch=worlds)
        agents = a
        steps = 6
    one = mk()
    many = mk()
    if agents is not None:
        one.upload(agents)
        many.upload(agents)
    one.step(steps)
    for _ in range(steps):
        many.step(1)
    if kind == "dynamic":
        r1, r2 = one.records(), many.records()
        for x, y, name in zip(r1, r2, ("id", "pos", "vel", "viewdir")):
            np.testing.assert_array_equal(x, y, err_msg=name)
    else:
        d1, d2 = one.download(), many.download()
        for name in ("pos", "vel", "viewdir"):
            np.testing.assert_array_equal(d1[name], d2[name], err_msg=name)
    one.close()
    many.close()
