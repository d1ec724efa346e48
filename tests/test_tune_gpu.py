"""Measurement and the exhaustive autotuner on the GPU: median-of-k timing, the
watchdog, the resumable results file and its best configuration."""
import csv

import pytest

import otlib
import ottune

pytestmark = pytest.mark.gpu


def test_measure_and_watchdog(otlib):
    p = otlib.make_params("jacobi-2d", nx=256, ny=256, problem_size=256, steps=8, T=2, x=32, y=4, e=2)
    d = otlib.derive(p)
    u = otlib.fill_uniform(d, 1, 0.0, 1.0)
    A, B = otlib.DeviceBuffer.from_array(u), otlib.DeviceBuffer.from_array(u)
    t, regs, to = ottune.measure(p, [A], [B], warmup=1, reps=5)
    assert 0 < t < 0.1 and regs > 0 and not to
    t2, _, to2 = ottune.measure(p, [A], [B], warmup=1, reps=5, time_limit_s=1e-9)
    assert to2 and t2 > 1e-9
    A.free()
    B.free()


def test_exhaustive_csv_resume_and_best(otlib, tmp_path):
    base = otlib.make_params("jacobi-2d", nx=96, ny=80, problem_size=96, steps=6, T=1, x=16, y=4, e=1)
    n_space = len(ottune.space(base))
    path = str(tmp_path / "sweep.csv")
    best, best_s, measured = ottune.exhaustive(base, path, warmup=1, reps=2)
    assert measured == n_space
    rows = list(csv.DictReader(open(path)))
    assert len(rows) == n_space
    assert {"Block Size X", "Block Size Y", "Time Tile Size", "Elements/Thread", "EventElapsed",
            "Register Usage"} <= set(rows[0])
    ok = [r for r in rows if r["Status"] == "ok"]
    assert len(ok) == n_space
    fastest = min(ok, key=lambda r: float(r["EventElapsed"]))
    assert best_s == pytest.approx(float(fastest["EventElapsed"]), rel=1e-6)
    assert (best.time_tile, best.block_x, best.block_y, best.elems_per_thread) == (
        int(fastest["Time Tile Size"]), int(fastest["Block Size X"]), int(fastest["Block Size Y"]),
        int(fastest["Elements/Thread"]))
    # resume: nothing left to measure, same best read back from the file
    best2, best_s2, measured2 = ottune.exhaustive(base, path, warmup=1, reps=2)
    assert measured2 == 0 and best_s2 == pytest.approx(best_s, rel=1e-6)
    assert len(list(csv.DictReader(open(path)))) == n_space
