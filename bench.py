"""Throughput of the per-step agent loop on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload mixed|bidirectional|sweep|circle|calibration]

A step is one WorldBase::DoStep (neighbour grid rebuild, neighbour search, new
velocity, contact forces, integration) over every agent, on synthetic agents already
resident in HBM. Default workload: the large mixed crowd, 1M agents in a 2000 x 2000 m
toric world (40% PowerLaw + sampling, 30% SocialForces + gradient with one gradient step
per frame, 30% ORCA, range 4 m, at most 12 neighbours, dt 0.05 s). The timed steps
replay the library's captured hipGraph of one step (umans_set_graph; --no-graph
launches them one by one). Prints one JSON line with the aggregate agent-steps/s, the
roofline of the dominant kernels (the class kernels of the step phase, timed with HIP
events on the library's stream over further eager steps) and the reference's own CPU
path timed on the host cores on a bounded sample of the same workload. The default
line adds three legs: the dense bidirectional flow (100k agents, Karamouzas 16 x 8 +
base), the batched sweep (512 worlds) and the adjoint of the calibration crowd (10k
agents, 100 steps), each with its roofline or check and its CPU baseline.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from umans_amd import roofline as RL  # noqa: E402
from umans_amd import scenarios as S  # noqa: E402


def workload(name):
    """(scenario, algorithmic FP32 operations per agent-step)."""
    if name == "mixed":
        return S.mixed(), RL.flops_mixed()
    if name == "bidirectional":
        return S.bidirectional(), RL.flops_karamouzas_sampling()
    if name == "circle":
        return S.circle(50), RL.flops_orca(10)
    if name == "sweep":
        return S.sweep(), RL.flops_powerlaw_sampling(12.6, 72)
    raise SystemExit(f"unknown workload {name}")


# ---------------------------------------------------------------- CPU baseline
def host_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


class _Quiet:
    """C-level stdout silenced (the reference reports its world on stdout; this
    program's stdout carries only the JSON line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.dn = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.dn, 1)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.dn)
        os.close(self.saved)


def _reference_available():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_py as R
    return R if R.available() else None


def _ref_sim(R, world_type, width, height, dt, policies, agents, threads):
    from umans_amd.registry import to_reference_xml_from_packed
    pols = [to_reference_xml_from_packed(p) for p in policies]
    with _Quiet():
        return R.RefSim(world_type, width, height, dt, pols, agents, threads=threads)


def cpu_baseline(sc, steps, threads):
    """The compiled reference (oracle/_ref, built from the reference's own sources; it
    travels with the tree) on the host cores: one CrowdSimulator on the same agents,
    one untimed step, then `steps` steps timed by the reference's own clock. Without it,
    the oracle's FP32 restatement (OpenMP)."""
    R = _reference_available()
    if R is not None:
        sim = _ref_sim(R, sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.agents, threads)
        with _Quiet():
            sim.step(1)
            t = sim.time_steps(steps)
        sim.close()
        kind = "reference"
    else:
        import oracle_py as O
        wd = O.make_world(sc.world_type, sc.width, sc.height, sc.dt, sc.max_neighbors)
        a = dict(sc.agents)
        t0 = time.perf_counter()
        for _ in range(steps):
            o = O.step(wd, sc.policies, a, precision=0, nthreads=threads)
            a["pos"], a["vel"], a["viewdir"] = o["pos"], o["vel"], o["viewdir"]
        t = time.perf_counter() - t0
        kind = "port"
    out = {"value": sc.n * steps / t, "unit": "agent-steps/s", "cores": threads, "kind": kind,
           "sample": f"{sc.n} agents x {steps} steps of the same workload ({t:.1f} s; the reference lists every "
                     f"neighbour in range, it has no cap)"}
    out.update(host_info())
    return out


def cpu_baseline_batch(sw, steps, threads, stride):
    """The reference has one world per CrowdSimulator: every stride-th world of the sweep
    as its own simulator, run one after another, each with one untimed step then `steps`
    timed steps; sum of N_s * steps over the sum of the times."""
    R = _reference_available()
    if R is None:
        return None
    import numpy as np
    starts = np.concatenate([[0], np.cumsum(sw.sizes)])
    agent_steps, t = 0, 0.0
    picked = list(range(0, len(sw.worlds), stride))
    for k in picked:
        sl = slice(starts[k], starts[k + 1])
        ak = {key: v[sl] for key, v in sw.agents.items() if key != "world"}
        ak["policy"] = np.zeros(sw.sizes[k], np.int32)
        wt, W, H, dt = sw.worlds[k]
        sim = _ref_sim(R, wt, W, H, dt, [sw.policies[k]], ak, threads)
        with _Quiet():
            sim.step(1)
            t += sim.time_steps(steps)
        sim.close()
        agent_steps += int(sw.sizes[k]) * steps
    out = {"value": agent_steps / t, "unit": "agent-steps/s", "cores": threads, "kind": "reference",
           "sample": f"{len(picked)} of {len(sw.worlds)} worlds (every {stride}th, {agent_steps // steps} agents) x "
                     f"{steps} steps, one simulator after another ({t:.1f} s)"}
    out.update(host_info())
    return out


def copy_probe(mib):
    """hipMemcpy device-to-device of `mib` MiB, three times, through the HIP runtime."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))
    return 0


# ---------------------------------------------------------------- measurements
def measured_traffic(profile, kernel="k_step"):
    """HBM bytes per step of `kernel` from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by the copy probe), or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", profile)):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 4 and cells[1] == kernel:
                return float(cells[4]) * 1e6
    except (OSError, ValueError):
        pass
    return None


# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD per 4 clocks
# (16 lanes wide, packed FP32 not counted) at 2.4 GHz, in G wave-instructions per second
VALU_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4


def measured_valu_per_step(profile):
    """VALU busy quad-cycles of one step (sum over the step kernels of waves x
    SQ_ACTIVE_INST_VALU per wave) from a committed rocprofv3 SQ summary, or None."""
    total = 0.0
    try:
        for line in open(os.path.join(ROOT, "profiles", profile)):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 12 and cells[1].startswith("k_step"):
                total += float(cells[2]) * float(cells[11])
    except (OSError, ValueError):
        return None
    return total or None


def roofline(n, step_s, flops_per_agent, traffic=None, valu_quads=None, kernel="k_step"):
    """The step kernels against both roofs that could bound them: HBM (84 algorithmic
    bytes per agent-step) and FP32 VALU (the model's operations per agent-step); `bound`
    is the one with the larger fraction."""
    gbs = RL.AGENT_STEP_BYTES * n / step_s / 1e9
    tfl = flops_per_agent * n / step_s / 1e12
    hbm = {"achieved": gbs, "peak": RL.HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / RL.HBM_PEAK_GBS,
           "bytes_per_agent": RL.AGENT_STEP_BYTES, "step_kernel_bytes_per_agent": RL.STEP_BYTES}
    valu = {"achieved": tfl, "peak": RL.FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tfl / RL.FP32_PEAK_TFLOPS,
            "flops_per_agent": flops_per_agent}
    top = valu if valu["frac"] >= hbm["frac"] else hbm
    out = {"bound": "valu" if top is valu else "hbm", "kernel": kernel, "achieved": top["achieved"],
           "peak": top["peak"], "unit": top["unit"], "frac": top["frac"], "traffic": traffic,
           "avg_launch_ms": step_s * 1e3, "hbm": hbm, "valu": valu}
    if valu_quads:
        g = valu_quads / step_s / 1e9
        out["valu_issue"] = {"achieved": g, "peak": VALU_ISSUE_PEAK_G, "unit": "G wave-instructions/s",
                             "frac": g / VALU_ISSUE_PEAK_G, "quad_cycles_per_step": valu_quads}
    return out


def gpu_run(w, steps, warmup, graph, phase_steps):
    """Timed steps (graph replays or eager launches), then eager steps with HIP events
    around the grid phase and the step kernels. Returns (seconds, profile)."""
    w.set_graph(graph)
    w.step(warmup)
    w.synchronize()
    t0 = time.perf_counter()
    w.step(steps)
    w.synchronize()
    t = time.perf_counter() - t0
    w.set_graph(False)
    w.set_profiling(True)
    w.step(max(phase_steps, 1))
    w.synchronize()
    prof = w.profile()
    w.set_profiling(False)
    return t, prof


def bidirectional_leg(args):
    from umans_amd import gpu
    sc, flops = workload("bidirectional")
    w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    w.upload(sc.agents)
    tuning = w.autotune(trial_steps=3)
    steps = 20
    t, prof = gpu_run(w, steps, 2, True, 5)
    w.close()
    step_s = prof["step_ms"] / 1e3 / max(prof["steps"], 1)
    out = {"metric": "agent-steps/s", "value": sc.n * steps / t, "unit": "agent-steps/s",
           "workload": sc.name, "agents": sc.n, "steps": steps, "ms_per_step": t * 1e3 / steps,
           "config": {"world": "toric", "size_m": sc.width, "dt": sc.dt, "max_neighbors": sc.max_neighbors,
                      "policies": "Karamouzas, regular sampling 16 angles x 8 speeds + base (129 candidates), range 5 m"},
           "tuning": tuning, "graph": True,
           "roofline": roofline(sc.n, step_s, flops,
                                traffic=measured_traffic("k_step_coop_bidirectional100k_hbm_counters.md",
                                                         "k_step_coop"), kernel="k_step_coop")}
    if not args.no_cpu_baseline:
        print("# bidirectional: timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps_bidirectional, args.cpu_threads)
        out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def descent_leg(args, tuning):
    """The 1M-agent mixed crowd with its SocialForces class descending 20 iterations of
    step 0.1 per frame (the gradient method's descent extension, umans_policy::
    gradient_iterations; the reference has only the single step, so there is no CPU
    baseline), at the launch tuning of the headline run."""
    from umans_amd import gpu
    sc = S.mixed()
    pols = [dict(p) for p in sc.policies]
    pols[1].update(gradient_iterations=20, gradient_step=0.1)
    w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, pols, sc.max_neighbors)
    w.upload(sc.agents)
    if tuning:
        w.set_tuning(tuning["cell_factor"], tuning["block_size"], list_capacity=tuning.get("list_capacity", 0),
                     class_block=tuning.get("class_block") or (0, 0, 0, 0, 0))
    steps = 20
    t, prof = gpu_run(w, steps, 3, True, 5)
    w.close()
    return {"metric": "agent-steps/s at 1M agents", "value": sc.n * steps / t, "unit": "agent-steps/s",
            "workload": sc.name, "steps": steps, "ms_per_step": t * 1e3 / steps,
            "phases_ms": {"grid": prof["grid_ms"] / max(prof["steps"], 1),
                          "step": prof["step_ms"] / max(prof["steps"], 1)},
            "config": {"policies": "as the headline, with the SocialForces class descending 20 iterations of "
                                   "step 0.1 per frame (GradientIterations=20, GradientStepSize=0.1)"}}


def batched_leg(args):
    from umans_amd import gpu
    sw, flops = workload("sweep")
    w = gpu.World(0, 0, 0, 0.1, sw.policies, max_neighbors=sw.max_neighbors, batch=sw.worlds)
    w.upload(sw.agents)
    tuning = w.autotune(trial_steps=6)
    steps = 100
    t, prof = gpu_run(w, steps, 3, True, 10)
    st = w.stats()
    w.close()
    step_s = prof["step_ms"] / 1e3 / max(prof["steps"], 1)
    out = {"metric": "batched agent-steps/s", "value": sw.n * steps / t, "unit": "agent-steps/s",
           "worlds": len(sw.worlds), "agents": sw.n, "steps": steps, "ms_per_step": t * 1e3 / steps,
           "config": {"world": "infinite", "dt": 0.1, "max_neighbors": sw.max_neighbors,
                      "policies": "per world PowerLaw (k~U[1,3], tau0~U[1,5]) + GoalReachingForce, sampling 12x6",
                      "parallelism": "single GPU, all worlds in each launch, one captured graph per step"},
           "tuning": tuning, "graph": True, "max_list": st["max_list"],
           "roofline": roofline(sw.n, step_s, flops, traffic=measured_traffic("k_step_sweep512_hbm_counters.md"),
                                valu_quads=measured_valu_per_step("k_step_sweep512_sq_counters.md"))}
    if not args.no_cpu_baseline:
        print("# batched: timing the CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline_batch(sw, args.cpu_steps_batched, args.cpu_threads, args.cpu_world_stride)
        if cb:
            cb["gpu_over_cpu"] = out["value"] / cb["value"]
            out["cpu_baseline"] = cb
    return out


def adjoint_leg(args, steps=100, ck=10):
    """Adjoint calibration crowd: 10k agents, 100 x 100 m toric world, SocialForces +
    GoalReachingForce (gradient method), dt 0.1 s, 100 steps. Targets from (A, B) =
    (2000, 0.08); gradients at (1500, 0.12) w.r.t. (A, B) and the 10k initial
    velocities, checkpoints every 10 steps. Reports the backward time per step against
    the plain step and, unless --no-adjoint-check, the gradients' relative error against
    the float64 reverse sweep along the same trajectory (tests/adjoint_ref.py, the
    checker)."""
    import numpy as np
    from umans_amd import adjoint as ADJ, gpu
    sc = S.calibration()
    w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    tg = ADJ.rollout_targets(w, sc, 2000.0, 0.08, steps)
    A0, B0 = 1500.0, 0.12

    def forward():
        w.set_policies([S.with_sf_params(p, A0, B0) for p in sc.policies])
        w.upload(sc.agents)
        return w.adjoint_forward(tg, steps, ck)

    forward()                                          # warm-up
    w.adjoint_backward(1)
    w.upload(sc.agents)
    w.synchronize()
    t0 = time.perf_counter()
    w.step(steps)
    w.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = forward()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    dp, _, dv = w.adjoint_backward(1)
    t_bwd = time.perf_counter() - t0
    clamps = w.adjoint_clamps(steps)
    out = {"metric": "backward/forward time", "value": t_bwd / t_plain, "unit": "ratio",
           "workload": sc.name, "agents": sc.n, "steps": steps, "checkpoint_every": ck,
           "plain_ms_per_step": t_plain * 1e3 / steps, "adjoint_forward_ms_per_step": t_fwd * 1e3 / steps,
           "backward_ms_per_step": t_bwd * 1e3 / steps, "loss": loss,
           "dL_dA": float(dp[0, 0]), "dL_dB": float(dp[0, 1]), "dL_dv0_norm": float(np.linalg.norm(dv))}
    if not args.no_adjoint_check:
        print("# adjoint: float64 reverse sweep along the same trajectory (the checker)", file=sys.stderr, flush=True)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import adjoint_ref as AR
        import torch
        states, lists = ADJ.record_rollout(w, sc, A0, B0, steps)
        c = AR.consts_from(sc)
        rec = [AR.lists_from_record(cnt, ids, seg, c["W"], c["H"]) for cnt, ids, seg in lists]
        dev = "cuda" if torch.cuda.is_available() else "cpu"
        _, g0, g1, _, gv = AR.replay_gradient(states, rec, tg, A0, float(np.float32(B0)), c, device=dev, clamps=clamps)
        out["check"] = {"reference": "float64 reverse sweep along the FP32 trajectory, same lists and branches",
                        "rel_err_dA": float(abs(dp[0, 0] - g0) / abs(g0)),
                        "rel_err_dB": float(abs(dp[0, 1] - g1) / abs(g1)),
                        "rel_err_dv0_l2": float(np.linalg.norm(dv - gv) / np.linalg.norm(gv)),
                        "fp64_dL_dA": float(g0), "fp64_dL_dB": float(g1)}
    w.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed")
    ap.add_argument("--cpu-steps", type=int, default=15, help="timed steps of the CPU baseline (1M agents: ~10 s)")
    ap.add_argument("--cpu-steps-bidirectional", type=int, default=5)
    ap.add_argument("--cpu-steps-batched", type=int, default=20)
    ap.add_argument("--cpu-world-stride", type=int, default=8, help="batched baseline: every n-th world")
    ap.add_argument("--cpu-threads", type=int, default=os.cpu_count() or 1, help="default: every host core")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="", help="CELL_FACTOR,BLOCK[,B0:B1:B2:B3:B4]: fix the launch tuning "
                    "(per-class block sizes optional, 0 = BLOCK) instead of autotuning (profiling runs use "
                    "the tuning the timed run chose)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the bidirectional, batched-sweep and adjoint legs added to the default line")
    ap.add_argument("--no-adjoint-check", action="store_true", help="skip the adjoint's float64 check")
    ap.add_argument("--no-graph", action="store_true", help="time eagerly launched steps instead of "
                    "replays of the captured hipGraph of one step")
    ap.add_argument("--phase-steps", type=int, default=20,
                    help="eager steps timed per phase with HIP events after the timed run (roofline)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (3 times): the known-size "
                         "reference that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe_mib:
        return copy_probe(args.copy_probe_mib)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the multi-GPU path is a documented "
              "design, not built.", file=sys.stderr)
        return 2

    from umans_amd import gpu
    if args.workload == "calibration":
        print(json.dumps(adjoint_leg(args), default=float))
        return 0
    sc, flops_per_agent = workload(args.workload)
    batched = isinstance(sc, S.Batch)
    if batched:
        w = gpu.World(0, 0, 0, 0.1, sc.policies, max_neighbors=sc.max_neighbors, batch=sc.worlds)
    else:
        w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    w.upload(sc.agents)
    if args.tuning:
        parts = args.tuning.split(",")
        cf, bs = float(parts[0]), int(parts[1])
        cb = [int(x) for x in parts[2].split(":")] if len(parts) > 2 else [0] * 5
        w.set_tuning(cf, bs, class_block=cb)
        tuning = {"cell_factor": cf, "block_size": bs, "class_block": cb, "fixed": True}
    else:
        tuning = None if args.no_autotune else w.autotune(trial_steps=10)
    graph = not args.no_graph
    t, prof = gpu_run(w, args.steps, args.warmup, graph, args.phase_steps)
    st = w.stats()
    w.close()

    n = sc.n
    mixed = sc.name.startswith("mixed")
    value = n * args.steps / t
    step_s = prof["step_ms"] / 1e3 / max(prof["steps"], 1)
    grid_s = prof["grid_ms"] / 1e3 / max(prof["steps"], 1)
    prof_names = {"mixed": ("k_step_mixed1M_hbm_counters.md", "k_step_mixed1M_sq_counters.md"),
                  "bidirectional": ("k_step_coop_bidirectional100k_hbm_counters.md", None),
                  "sweep": ("k_step_sweep512_hbm_counters.md", "k_step_sweep512_sq_counters.md")}
    pf = prof_names.get(args.workload, (None, None))
    out = {
        "metric": ("agent-steps/s at 1M agents" if mixed else "batched agent-steps/s" if batched else "agent-steps/s"),
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": ({"workload": sc.name, "agents": n, "worlds": len(sc.worlds), "world": "infinite",
                    "dt": 0.1, "max_neighbors": sc.max_neighbors,
                    "policies": "per world PowerLaw(k~U[1,3], tau0~U[1,5]) + GoalReachingForce, sampling 12x6",
                    "parallelism": "single GPU, all worlds in each launch"} if batched else
                   {"workload": sc.name, "agents": n, "world": "toric" if sc.world_type == 1 else "infinite",
                    "size_m": sc.width, "dt": sc.dt, "max_neighbors": sc.max_neighbors,
                    "policies": ("40% PowerLaw+sampling 12x6, 30% SocialForces+gradient (one gradient step per "
                                 "frame, the reference's GradientMethod), 30% ORCA") if mixed else sc.name,
                    "parallelism": "single GPU"}),
        "tuning": tuning, "graph": graph,
        "roofline": roofline(n, step_s, flops_per_agent, traffic=measured_traffic(pf[0]) if pf[0] else None,
                             valu_quads=measured_valu_per_step(pf[1]) if pf[1] else None),
        "phases_ms": {"grid": grid_s * 1e3, "step": step_s * 1e3},
        "neighbor_overflow_agents": st["overflow_agents"],
    }
    if not args.no_cpu_baseline:
        print(f"# GPU leg done ({value:.3e} agent-steps/s); timing the CPU baseline", file=sys.stderr, flush=True)
        if batched:
            cb = cpu_baseline_batch(sc, args.cpu_steps_batched, args.cpu_threads, args.cpu_world_stride)
        else:
            cb = cpu_baseline(sc, args.cpu_steps, args.cpu_threads)
        if cb:
            cb["gpu_over_cpu"] = value / cb["value"]
            out["cpu_baseline"] = cb
    if not args.no_extras and mixed:
        print("# timing the bidirectional, batched-sweep and adjoint legs", file=sys.stderr, flush=True)
        out["descent"] = descent_leg(args, tuning)
        out["bidirectional"] = bidirectional_leg(args)
        out["batched"] = batched_leg(args)
        out["adjoint"] = adjoint_leg(args)
    print(json.dumps(out, default=float))
    return 0


if __name__ == "__main__":
    sys.exit(main())
