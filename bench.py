"""Throughput of the per-step agent loop on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload mixed|bidirectional|circle]

A step is one WorldBase::DoStep (neighbour grid rebuild, neighbour search, new
velocity, contact forces, integration) over every agent, on synthetic agents already
resident in HBM. Default workload: the large mixed crowd, 1M agents in a 2000 x 2000 m
toric world (40% PowerLaw + sampling, 30% SocialForces + gradient, 30% ORCA, range 4 m,
at most 12 neighbours, dt 0.05 s). The timed steps replay the library's captured
hipGraph of one step (umans_set_graph; --no-graph launches them one by one). Prints one
JSON line with the aggregate agent-steps/s, the roofline of the dominant kernel (the
class kernels of the step phase, timed with HIP events on the library's stream over
further eager steps) and the reference's own CPU path timed on the host cores on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from umans_amd import roofline as RL  # noqa: E402
from umans_amd import scenarios as S  # noqa: E402


def workload(name):
    if name == "mixed":
        return S.mixed(), RL.flops_mixed()
    if name == "bidirectional":
        return S.bidirectional(), RL.flops_karamouzas_sampling()
    if name == "circle":
        return S.circle(50), RL.flops_orca(10)
    if name == "sweep":
        return S.sweep(), RL.flops_powerlaw_sampling(30, 72)
    if name == "calibration":
        return None, None
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(sc, steps, threads):
    """Time the CPU path on the host cores: the compiled reference (oracle/_ref) when it
    is present, else the oracle's FP32 restatement (OpenMP)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_py as R
    from umans_amd.registry import to_reference_xml_policy
    if R.available() and sc.name.startswith("mixed"):
        pols = [to_reference_xml_policy("sampling", [("PowerLaw", 1.0, {"k": 1.5, "tau0": 3.0, "range": 4.0}),
                                                     ("GoalReachingForce", 1.0, {})],
                                        sampling={"SamplingType": "regular", "SamplingBase": "zero",
                                                  "SamplingBaseDirection": "preferred velocity",
                                                  "SamplingAngle": 360, "SpeedSamples": 6, "AngleSamples": 12,
                                                  "SamplingRadius": "maximum speed"}, relaxation_time=0.5),
                to_reference_xml_policy("gradient", [("SocialForcesAvoidance", 1.0, {"V0": 2000.0, "sigma": 0.08,
                                                                                    "range": 4.0, "agentMass": 80.0}),
                                                     ("GoalReachingForce", 1.0, {})], relaxation_time=0.5),
                to_reference_xml_policy("global", [("ORCA", 1.0, {"timeHorizon": 3.0, "range": 4.0})])]
        # the reference reports its world on stdout; keep this program's stdout to the JSON line
        sys.stdout.flush()
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            sim = R.RefSim(sc.world_type, sc.width, sc.height, sc.dt, pols, sc.agents, threads=threads)
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
            os.close(saved)
        sim.step(1)   # first step outside the timing, as the warm-up of the GPU leg
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
    return {"value": sc.n * steps / t, "unit": "agent-steps/s", "cores": threads, "kind": kind,
            "sample": f"{sc.n} agents x {steps} steps of the same workload ({t:.1f} s)"}


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


def adjoint_leg(steps=100, ck=10):
    """Adjoint calibration crowd: 10k agents, 100 x 100 m toric world, SocialForces +
    GoalReachingForce (gradient method), dt 0.1 s, 100 steps. Targets from (A, B) =
    (2000, 0.08); gradients at (1500, 0.12) w.r.t. (A, B) and the 10k initial
    velocities, checkpoints every 10 steps. Reports forward and backward time per step
    against the plain step, and central differences (h = 1e-3 relative, FP32 GPU
    forward) next to the adjoint."""
    import numpy as np
    from umans_amd import adjoint as ADJ, gpu
    sc = S.calibration()
    w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
    tg = ADJ.rollout_targets(w, sc, 2000.0, 0.08, steps)
    A0, B0 = 1500.0, 0.12

    def loss_at(A, B):
        w.set_policies([S.with_sf_params(p, A, B) for p in sc.policies])
        w.upload(sc.agents)
        return w.adjoint_forward(tg, steps, ck)

    loss_at(A0, B0)                                   # warm-up
    w.upload(sc.agents)
    w.synchronize()
    t0 = time.perf_counter()
    w.step(steps)
    w.synchronize()
    t_plain = time.perf_counter() - t0
    w.set_policies([S.with_sf_params(p, A0, B0) for p in sc.policies])
    w.upload(sc.agents)
    t0 = time.perf_counter()
    loss = w.adjoint_forward(tg, steps, ck)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    dp, _, dv = w.adjoint_backward(1)
    t_bwd = time.perf_counter() - t0
    fd = []
    for i, (x, h) in enumerate(((A0, 1e-3 * A0), (B0, 1e-3 * B0))):
        a = [A0, B0]
        a[i] = x + h
        lp = loss_at(*a)
        a[i] = x - h
        lm = loss_at(*a)
        fd.append((lp - lm) / (2 * h))
    w.close()
    return {"metric": "backward/forward time", "value": t_bwd / t_plain, "unit": "ratio",
            "workload": sc.name, "agents": sc.n, "steps": steps, "checkpoint_every": ck,
            "plain_ms_per_step": t_plain * 1e3 / steps, "adjoint_forward_ms_per_step": t_fwd * 1e3 / steps,
            "backward_ms_per_step": t_bwd * 1e3 / steps, "loss": loss,
            "dL_dA": float(dp[0, 0]), "dL_dB": float(dp[0, 1]), "dL_dv0_norm": float(np.linalg.norm(dv)),
            "fd_h1e-3_dA": fd[0], "fd_h1e-3_dB": fd[1]}


def measured_traffic(kernel="k_step"):
    """HBM bytes per launch of `kernel` on the mixed crowd from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by the copy probe), or None."""
    path = os.path.join(ROOT, "profiles", "k_step_mixed1M_hbm_counters.md")
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 4 and cells[1] == kernel:
                return float(cells[4]) * 1e6
    except OSError:
        pass
    return None


# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD per 4 clocks
# (16 lanes wide, packed FP32 not counted) at 2.4 GHz, in G wave-instructions per second
VALU_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4


def measured_valu_per_step():
    """VALU busy quad-cycles of one mixed-crowd step (sum over the class kernels of waves x
    SQ_ACTIVE_INST_VALU per wave) from the committed rocprofv3 SQ summary, or None."""
    path = os.path.join(ROOT, "profiles", "k_step_mixed1M_sq_counters.md")
    total = 0.0
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 12 and cells[1].startswith("k_step,"):
                total += float(cells[2]) * float(cells[11])
    except (OSError, ValueError):
        return None
    return total or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default="", help="CELL_FACTOR,BLOCK[,B0:B1:B2:B3:B4]: fix the launch tuning "
                    "(per-class block sizes optional, 0 = BLOCK) instead of autotuning (profiling runs use "
                    "the tuning the timed run chose)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched-sweep and adjoint legs added to the default line")
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
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "documented, not run.", file=sys.stderr)
        return 2

    from umans_amd import gpu
    if args.workload == "calibration":
        print(json.dumps(adjoint_leg()))
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
    # the timed run: replays of the captured step (or eager launches with --no-graph)
    w.set_graph(graph)
    w.step(args.warmup)
    w.synchronize()
    t0 = time.perf_counter()
    w.step(args.steps)
    w.synchronize()
    t = time.perf_counter() - t0
    st = w.stats()
    # per-phase timing for the roofline: eager steps with HIP events on the library's
    # stream around the grid phases and around the class kernels' fork and join
    w.set_graph(False)
    w.set_profiling(True)
    w.step(max(args.phase_steps, 1))
    w.synchronize()
    prof = w.profile()
    w.close()

    n = sc.n
    value = n * args.steps / t
    step_s = prof["step_ms"] / 1e3 / max(prof["steps"], 1)
    grid_s = prof["grid_ms"] / 1e3 / max(prof["steps"], 1)
    achieved = RL.STEP_BYTES * n / step_s / 1e9
    valu = flops_per_agent * n / step_s / 1e12
    # the bound the step kernels meet: VALU issue (quad-cycles per step from the counters)
    valu_issue = None
    qc = measured_valu_per_step() if sc.name.startswith("mixed") else None
    if qc:
        g = qc / step_s / 1e9
        valu_issue = {"achieved": g, "peak": VALU_ISSUE_PEAK_G, "unit": "G wave-instructions/s",
                      "frac": g / VALU_ISSUE_PEAK_G, "quad_cycles_per_step": qc}
    out = {
        "metric": ("agent-steps/s at 1M agents" if sc.name.startswith("mixed") else
                   "batched agent-steps/s" if batched else "agent-steps/s"),
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
                    "policies": "40% PowerLaw+sampling 12x6, 30% SocialForces+gradient, 30% ORCA"
                    if sc.name.startswith("mixed") else sc.name, "parallelism": "single GPU"}),
        "tuning": tuning, "graph": graph,
        "roofline": {"bound": "hbm", "kernel": "k_step", "achieved": achieved, "peak": RL.HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / RL.HBM_PEAK_GBS,
                     "traffic": measured_traffic() if sc.name.startswith("mixed") else None,
                     "bytes_per_agent": RL.STEP_BYTES, "avg_launch_ms": step_s * 1e3,
                     "valu": {"achieved": valu, "peak": RL.FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": valu / RL.FP32_PEAK_TFLOPS, "flops_per_agent": flops_per_agent},
                     "valu_issue": valu_issue},
        "phases_ms": {"grid": grid_s * 1e3, "step": step_s * 1e3},
        "neighbor_overflow_agents": st["overflow_agents"],
    }
    if not args.no_cpu_baseline and not batched:
        print(f"# GPU leg done ({value:.3e} agent-steps/s); timing the CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps, args.cpu_threads)
        out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    if not args.no_extras and sc.name.startswith("mixed"):
        print("# timing the batched sweep and the adjoint legs", file=sys.stderr, flush=True)
        sw, _ = workload("sweep")
        wb = gpu.World(0, 0, 0, 0.1, sw.policies, max_neighbors=sw.max_neighbors, batch=sw.worlds)
        wb.upload(sw.agents)
        tb = wb.autotune(trial_steps=6)
        wb.set_graph(True)
        wb.step(3)
        wb.synchronize()
        steps_b = 100
        t0 = time.perf_counter()
        wb.step(steps_b)
        wb.synchronize()
        tbt = time.perf_counter() - t0
        wb.close()
        out["batched"] = {"metric": "batched agent-steps/s", "value": sw.n * steps_b / tbt, "unit": "agent-steps/s",
                          "worlds": len(sw.worlds), "agents": sw.n, "steps": steps_b,
                          "ms_per_step": tbt * 1e3 / steps_b, "graph": True, "tuning": tb}
        out["adjoint"] = adjoint_leg()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
