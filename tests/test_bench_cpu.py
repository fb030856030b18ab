"""bench.py's CPU-side pieces: the figures its roofline takes from the committed
rocprofv3 summaries parse (a reformatted table would silently drop them), the roofline
arithmetic uses the frozen per-agent-step constants and reports the larger fraction as
the bound, and a run asked for more than one GPU exits with a message before touching a
GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from umans_amd import roofline as RL  # noqa: E402


@pytest.mark.parametrize("profile,kernel", [("k_step_mixed1M_hbm_counters.md", "k_step"),
                                            ("k_step_sweep512_hbm_counters.md", "k_step"),
                                            ("k_step_coop_bidirectional100k_hbm_counters.md", "k_step_coop")])
def test_committed_traffic_parses(profile, kernel):
    t = bench.measured_traffic(profile, kernel)
    assert t is not None and 1e6 < t < 1e10


@pytest.mark.parametrize("profile", ["k_step_mixed1M_sq_counters.md", "k_step_sweep512_sq_counters.md"])
def test_committed_valu_counts_parse(profile):
    q = bench.measured_valu_per_step(profile)
    assert q is not None and q > 1e6


def test_roofline_arithmetic():
    n, step_s, flops = 1_000_000, 0.4e-3, 1785.4
    r = bench.roofline(n, step_s, flops, traffic=443e6, valu_quads=168e6)
    assert r["hbm"]["achieved"] == pytest.approx(RL.AGENT_STEP_BYTES * n / step_s / 1e9)
    assert r["valu"]["achieved"] == pytest.approx(flops * n / step_s / 1e12)
    top = max(r["hbm"]["frac"], r["valu"]["frac"])
    assert r["frac"] == pytest.approx(top)
    assert r["bound"] == ("valu" if r["valu"]["frac"] >= r["hbm"]["frac"] else "hbm")
    assert r["traffic"] == 443e6


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""
