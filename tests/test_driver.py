"""The ot-<stencil> drivers: reference CLI and stdout wire format.

CPU tests cover option handling; GPU tests run the drivers with --verify and
check their derived keys against the reference's recorded sweep rows."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ics-12-overlapped-tiling_amd", "bin")


def run_driver(name, *args, timeout=120):
    r = subprocess.run([os.path.join(BIN, f"ot-{name}"), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout)
    keys = {}
    for line in r.stdout.splitlines():
        if ": " in line:
            k, v = line.split(": ", 1)
            keys[k] = v
    return r.returncode, keys, r


def test_driver_help_exits_1():
    rc, _, r = run_driver("jacobi-2d", "--help")
    assert rc == 1 and "--time-tile-size" in r.stderr


def test_driver_rejects_unknown_option():
    rc, _, r = run_driver("jacobi-3d", "--bogus", "1")
    assert rc == 1


@pytest.mark.parametrize("args", [["-n", "abc"], ["-n", "12abc"], ["-t", ""], ["-s", "3.5"], ["-x", "99999999999"],
                                  ["--alpha", "fast"], ["--seed", "-1"], ["-c", "maybe"],
                                  ["--time-only", "gpu"]])
def test_driver_malformed_value_prints_usage_and_exits_1(args):
    rc, keys, r = run_driver("jacobi-2d", *args)
    assert rc == 1 and "Usage:" in r.stderr and not keys


def test_time_only_and_dump_clocks_exist_only_for_the_2d_lds_drivers():
    # refused before any device work: the 3D drivers have no --time-only, and
    # clock records only for jacobi-3d tiles of the wave-pipelined kernel (the
    # reference geometry's default tile is not one)
    rc, keys, r = run_driver("jacobi-3d", "-n", 32, "--time-only", "gmem")
    assert rc == 1 and "jacobi-2d and gradient-2d" in r.stdout
    rc, keys, r = run_driver("jacobi-3d", "-n", 32, "-c", "1")
    assert rc == 1 and "jacobi-2d and gradient-2d" in r.stdout
    rc, keys, r = run_driver("poisson-3d-rhs", "-n", 64, "-s", 2, "-x", 60, "-y", 4, "-z", 8, "--geometry", "tile",
                             "-c", "1")
    assert rc == 1 and "jacobi-3d tiles of the wave-pipelined kernel" in r.stdout


def test_every_stencil_has_a_driver(otlib):
    for name in otlib.STENCILS:
        assert os.path.exists(os.path.join(BIN, f"ot-{name}")), name


@pytest.mark.gpu
@pytest.mark.parametrize("name,args", [
    ("jacobi-2d", ["-n", 1024, "-t", 64, "-s", 4, "-x", 16, "-y", 16]),
    ("jacobi-3d", ["-n", 64, "-t", 9, "-s", 2, "-x", 8, "-y", 8, "-z", 8]),
    ("jacobi-1d", ["-n", 5000, "-t", 17, "-s", 3, "-x", 64, "-e", 2]),
    ("jacobi-1d-5pt", ["-n", 3000, "-t", 9, "-s", 2, "-x", 64]),
    ("jacobi-1d-7pt", ["-n", 3000, "-t", 9, "-s", 2, "-x", 64]),
    ("poisson-2d", ["-n", 200, "-t", 8, "-s", 2, "-x", 16, "-y", 8]),
    ("poisson-3d", ["-n", 40, "-t", 4, "-s", 2, "-x", 8, "-y", 8, "-z", 8]),
    ("poisson-3d-rhs", ["-n", 40, "-t", 9, "-s", 3, "-x", 16, "-y", 16, "-z", 8, "--geometry", "tile"]),
    ("gradient-2d", ["-n", 300, "-t", 16, "-s", 4, "-x", 32, "-y", 8, "-e", 2]),
    ("gradient-3d", ["-n", 50, "-t", 8, "-s", 2, "-x", 8, "-y", 8, "-z", 8]),
    ("fdtd-2d", ["-n", 300, "-t", 20, "-s", 5, "-x", 32, "-y", 16]),
    ("rician-denoise-2d", ["-n", 300, "-t", 20, "-s", 2, "-x", 32, "-y", 8]),
    ("cdsc-tv-update-2d", ["-n", 300, "-t", 20, "-s", 4, "-x", 32, "-y", 16, "--graph"]),
    ("cs-3d", ["-n", 100, "-t", 9, "-s", 2, "-x", 60, "-y", 8, "-z", 8, "--geometry", "tile"]),
])
def test_driver_verify_passes(name, args):
    rc, keys, r = run_driver(name, *args, "-v")
    print(r.stdout[-600:], flush=True)
    assert rc == 0, r.stdout + r.stderr
    assert keys["Correctness"] == "PASSED"
    assert float(keys["Max Relative Error"]) <= 1e-5
    for k in ("EventElapsed", "Elapsed Time", "Actual GFlop/s", "Device GFlop/s", "Register Usage",
              "Num Blocks X", "Padded Size", "Real Size", "phase2_global_loads", "Dimensions"):
        assert k in keys, k


@pytest.mark.gpu
def test_driver_keys_match_reference_sweep_rows():
    """A few rows of the reference's jacobi-2d C2050 sweep: the driver prints the
    same derived keys for the same command line."""
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "geometry_c2050.json")))
    rows = [r for r in golden if r["stencil"] == "jacobi-2d"][:4] + \
           [r for r in golden if r["stencil"] == "jacobi-3d"][:3]
    for r in rows:
        args = ["-n", r["problem_size"], "-t", 2, "-s", r["Time Tile Size"], "-x", r["Block Size X"],
                "-y", r["Block Size Y"], "-e", r["Elements/Thread"]]
        if r["stencil"] == "jacobi-3d":
            args += ["-z", r["Block Size Z"]]
        rc, keys, out = run_driver(r["stencil"], *args, timeout=300)
        assert rc == 0, out.stdout + out.stderr
        for k in ("Num Blocks X", "Num Blocks Y", "Num Blocks Z", "Padded Size", "Shared Size",
                  "Global Loads/Block", "Shared Loads/Block"):
            if k in r:
                assert int(keys[k]) == int(r[k]), (k, r)


@pytest.mark.gpu
def test_save_and_load_plan(tmp_path):
    plan = tmp_path / "plan.txt"
    rc, k1, _ = run_driver("jacobi-2d", "-n", 500, "-t", 8, "-s", 2, "-w", plan)
    assert rc == 0 and plan.exists()
    rc, k2, _ = run_driver("jacobi-2d", "-f", plan, "-v")
    assert rc == 0 and k2["Correctness"] == "PASSED"
    assert k1["Num Blocks X"] == k2["Num Blocks X"] and k1["Time Tile Size"] == k2["Time Tile Size"]


def _clock_rows(stderr):
    lines = stderr.splitlines()
    i = lines.index("bx,by,warp,sm,clock0,clock1,clock2,clock3,clock4,")
    return [[int(v) for v in ln.rstrip(",").split(",")] for ln in lines[i + 1:] if ln.count(",") == 9]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["jacobi-2d", "gradient-2d"])
def test_driver_dump_clocks_and_time_only(name):
    """-c 1 prints the reference's ClockData CSV (jacobi-2d.cpp:957-997) for the
    waves of one compute unit, with stamps in phase order, and leaves the
    result unchanged (verify passes); every --time-only mode runs to the end."""
    base = ["-n", 700, "-t", 12, "-s", 4, "-x", 32, "-y", 8, "-e", 2]
    rc, keys, r = run_driver(name, *base, "-c", 1, "-v")
    assert rc == 0, r.stdout + r.stderr
    assert keys["Correctness"] == "PASSED"
    rows = _clock_rows(r.stderr)
    assert rows and all(row[3] == 1 for row in rows)
    for row in rows:
        c = row[4:]
        assert c[0] <= c[1] <= c[2] <= c[3] <= c[4], row
    for mode in ("gmem", "instr", "smem"):
        rc, keys, r = run_driver(name, *base, "--time-only", mode, "-v")
        assert rc == 0, (mode, r.stdout + r.stderr)
        assert float(keys["EventElapsed"]) > 0 and keys["Correctness"].startswith("SKIPPED")
