"""CSV trajectory output (CrowdSimulator::StartCSVOutput, crowdSimulator.cpp:37-51, and
CSVWriter, csvwriter.cpp:37-87).

The reference writes one file output_<id>.csv per agent; every step appends a row
"t,x,y" where t is the world time (double, WorldBase::time_ += float delta_time,
worldBase.cpp:122) and x, y the float position, all in the default stream format
(6 significant digits, %g). The CPU test pins that row format against the text the
reference's writer produced (fixture tests/golden/csv_Circle25-ORCA.npz); the GPU tests
check the library's files row for row against the positions the library itself
downloads after each step, and against the reference's trajectory."""
import os

import numpy as np
import pytest

from goldens import GOLDEN
from umans_amd import gpu

EXAMPLE = os.path.join(GOLDEN, "examples", "Circle25-ORCA.xml")
FIXTURE = os.path.join(GOLDEN, "csv_Circle25-ORCA.npz")


def csv_row(t, x, y):
    """One row as CSVWriter::Flush writes it (csvwriter.cpp:66): operator<< on a double
    and two floats with the default precision 6 and no floatfield, i.e. %g."""
    return "%g,%g,%g\n" % (t, float(np.float32(x)), float(np.float32(y)))


def world_times(dt, nsteps):
    t, out = 0.0, []
    for _ in range(nsteps):
        t += float(np.float32(dt))
        out.append(t)
    return out


def expected_text(times, positions):
    """positions: (steps, n, 2) float32 in row order -> one text per row."""
    n = positions.shape[1]
    return ["".join(csv_row(t, *positions[s, i]) for s, t in enumerate(times)) for i in range(n)]


def read_csv_dir(dirname, ids):
    return [open(os.path.join(dirname, f"output_{i}.csv")).read() for i in ids]


def test_row_format_matches_reference_writer():
    d = np.load(FIXTURE)
    sc, _ = gpu.load_scenario(EXAMPLE)
    steps = d["state_pos"].shape[0] - 1
    assert (d["state_id"] == d["id"][None]).all()
    want = expected_text(world_times(sc.dt, steps), d["state_pos"][1:])
    assert [str(x) for x in d["text"]] == want


def test_row_format_catches_precision_slips():
    """Positions printed with more digits (or times printed in full) would not match the
    reference's text."""
    d = np.load(FIXTURE)
    sc, _ = gpu.load_scenario(EXAMPLE)
    steps = d["state_pos"].shape[0] - 1
    times = world_times(sc.dt, steps)
    long_rows = ["".join("%.9g,%.9g,%.9g\n" % (t, *p) for t, p in zip(times, d["state_pos"][1:, i]))
                 for i in range(len(d["id"]))]
    assert [str(x) for x in d["text"]] != long_rows
    full_t = ["".join("%r,%g,%g\n" % (t, *p) for t, p in zip(times, d["state_pos"][1:, i]))
              for i in range(len(d["id"]))]
    assert [str(x) for x in d["text"]] != full_t


@pytest.mark.gpu
@pytest.mark.parametrize("flush_now", [False, True])
def test_gpu_csv_rows_match_downloads(tmp_path, flush_now):
    out = tmp_path / "traj" / "run"
    w = gpu.World.from_scenario_file(EXAMPLE)
    ids = w.scenario.agents["id"]
    w.csv_start(str(out), flush_immediately=flush_now)
    pos = []
    for s in range(6):
        w.step(1)
        pos.append(w.download()["pos"])
        if flush_now:
            rows = read_csv_dir(out, ids[:1])[0].count("\n")
            assert rows == s + 1
    w.step(4)                                   # several steps in one call go through the pinned slots
    w.csv_flush()
    w.csv_stop()
    ref = gpu.World.from_scenario_file(EXAMPLE)  # the same run without output, for steps 7..10
    ref.step(6)
    for s in range(4):
        ref.step(1)
        pos.append(ref.download()["pos"])
    pos = np.stack(pos)
    want = expected_text(world_times(w.scenario.dt, 10), pos)
    got = read_csv_dir(out, ids)
    assert got == want
    assert sorted(os.listdir(out)) == sorted(f"output_{i}.csv" for i in ids)


@pytest.mark.gpu
def test_gpu_csv_follows_reference_trajectory(tmp_path):
    """Over the first 10 steps of the ORCA circle the library's rows hold the reference's
    times exactly and its positions to 1e-3 m."""
    d = np.load(FIXTURE)
    w = gpu.World.from_scenario_file(EXAMPLE)
    w.csv_start(str(tmp_path))
    w.step(d["state_pos"].shape[0] - 1)
    w.csv_stop()
    got = read_csv_dir(tmp_path, d["id"])
    for g, r in zip(got, d["text"]):
        gt = np.array([[float(v) for v in ln.split(",")] for ln in g.splitlines()])
        rt = np.array([[float(v) for v in ln.split(",")] for ln in str(r).splitlines()])
        assert [ln.split(",")[0] for ln in g.splitlines()] == [ln.split(",")[0] for ln in str(r).splitlines()]
        np.testing.assert_allclose(gt[:, 1:], rt[:, 1:], atol=1e-3)


@pytest.mark.gpu
def test_gpu_csv_appends_and_stops(tmp_path):
    """Flush appends (std::ios::app); after csv_stop no further rows are written."""
    w = gpu.World.from_scenario_file(EXAMPLE)
    ids = w.scenario.agents["id"]
    w.csv_start(str(tmp_path))
    w.step(2)
    w.csv_flush()
    w.step(3)
    w.csv_stop()
    w.step(2)
    got = read_csv_dir(tmp_path, ids)
    assert all(g.count("\n") == 5 for g in got)
    times = [ln.split(",")[0] for ln in got[0].splitlines()]
    assert times == ["%g" % t for t in world_times(w.scenario.dt, 5)]
