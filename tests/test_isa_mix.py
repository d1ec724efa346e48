"""The wave-pipelined jacobi-3d kernel's steady-state loop, in the gfx950 ISA
of this build, issues nothing but the update itself: per unrolled block of 6
planes, 24 row evaluations per plane (stage 1 on footprint rows 1-10, stage 2
on 2-9, stage 3 on 3-8 at T=3 with 12 footprint rows), each 6 adds (2 of them
with a DPP x-neighbour operand) and 1 multiply (jacobi-3d.cpp:734-741: 0.143f
times a sum of seven taps), and 12 loads and 6 stores per plane. Compiles
csrc/k3w_j3d.hip with hipcc (no GPU needed, about a minute)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "ics-12-overlapped-tiling_amd", "isa_mix.py")


@pytest.mark.skipif(not shutil.which("/opt/rocm/bin/hipcc"), reason="hipcc not installed")
def test_j3d_steady_loop_is_the_update_only():
    r = subprocess.run([sys.executable, TOOL], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    counts = {}
    for line in r.stdout.splitlines()[1:]:
        n, op = line.split(None, 1)
        counts[op.strip()] = int(n)
    evals = 24 * 6
    assert counts.get("v_add_f32_e32", 0) == 4 * evals
    assert counts.get("v_add_f32_dpp (dpp)", 0) == 2 * evals
    assert counts.get("v_mul_f32_e32", 0) == evals
    assert sum(v for k, v in counts.items() if k.startswith("v_")) == 7 * evals
    assert counts.get("buffer_load_dword", 0) == 12 * 6
    assert counts.get("buffer_store_dword", 0) == 6 * 6
