"""Tuning host logic on the CPU: the configuration space, the feature vector and
the analytic model (restatement of scripts/run-model.py), pinned by a
hand-derived case and by the model's own phase decomposition."""
import math

import numpy as np
import pytest

import otlib
import ottune


def _p(name, n3, steps, T, x, y, z, e, **kw):
    return otlib.make_params(name, nx=n3[0], ny=n3[1], nz=n3[2], problem_size=n3[0], steps=steps, T=T, x=x, y=y,
                             z=z, e=e, **kw)


@pytest.mark.parametrize("name,n3", [("jacobi-1d", (1 << 16, 1, 1)), ("jacobi-2d", (1024, 1024, 1)),
                                     ("rician-denoise-2d", (300, 200, 1)), ("jacobi-3d", (256, 256, 256)),
                                     ("poisson-3d", (64, 64, 64))])
def test_space_is_valid_and_unique(otlib, name, n3):
    sp = ottune.space(_p(name, n3, 16, 1, 16, 4, 4, 1))
    assert len(sp) > 50
    keys = set()
    for q in sp:
        otlib.derive(q)  # raises if the generator's checks reject it
        keys.add((q.time_tile, q.block_x, q.block_y, q.block_z, q.elems_per_thread))
    assert len(keys) == len(sp)


def test_space_holds_both_3d_kernel_families(otlib):
    sp = ottune.space(_p("jacobi-3d", (256, 256, 256), 32, 1, 16, 4, 4, 1))
    keys = {(q.time_tile, q.block_x, q.block_y, q.block_z, q.elems_per_thread) for q in sp}
    assert (3, 58, 6, 20, 1) in keys  # wave-pipelined
    assert (2, 32, 4, 32, 4) in keys  # LDS
    wave = [q for q in sp if ottune.features(q)[23] == 1.0]
    assert 0 < len(wave) < len(sp)


def test_space_balances_wave_counts_over_z(otlib):
    """3D wave tiles also get the z-chunks that make the wave count just fill k
    waves per SIMD (k = 1..6) on 256 CUs x 4 SIMDs: J3D 256^3 with a 58x6 tile has
    5 x 43 = 215 waves per z-layer of chunks, so k = 3 gives 3 * 1024 // 215 = 14
    chunks of ceil(256 / 14) = 19 planes, k = 2 gives 9 chunks of 29."""
    sp = ottune.space(_p("jacobi-3d", (256, 256, 256), 32, 1, 16, 4, 4, 1))
    bz = {q.block_z for q in sp if (q.time_tile, q.block_x, q.block_y) == (3, 58, 6)}
    assert {8, 16, 20, 24, 32} <= bz and {19, 29, 64} <= bz
    for k in range(1, 7):
        chunks = k * 1024 // 215
        assert -(-256 // chunks) in bz


def test_features(otlib):
    p = _p("jacobi-3d", (256, 128, 64), 32, 3, 58, 6, 20, 1)
    f = ottune.features(p)
    assert f.shape == (28,)
    assert f[4] == 1.0 and f[:13].sum() == 1.0
    assert tuple(f[13:16]) == (8.0, 7.0, 6.0)
    assert f[16] == 1 and f[17] == 1 and f[18] == 3
    assert f[19] == pytest.approx(math.log2(58)) and f[20] == pytest.approx(math.log2(6))
    assert f[21] == pytest.approx(math.log2(20)) and f[22] == 0.0 and f[23] == 1.0
    assert tuple(f[24:27]) == (3.0, 7.0, 7.0)  # dims, flops per point, loads per point
    prior = ottune.analytic_seconds(p, 0) / (256 * 128 * 64 * 32)
    assert f[27] == pytest.approx(math.log(prior), rel=1e-6)
    g = ottune.features(_p("poisson-3d-rhs", (64, 64, 64), 8, 2, 32, 4, 16, 4))
    assert g[7] == 1.0 and g[17] == 2 and g[22] == 2.0 and g[23] == 0.0


def _arch(**kw):
    a = ottune.default_arch()
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def test_analytic_hand_derived_case(otlib):
    """Jacobi-2D 256^2, T=1, 64x4 tile, 4 steps, on a machine with unit issue
    costs and zero bandwidth / latency / launch terms and no L1. Per
    run-model.py: p2_glb = 5 (no hits), p2_shd = 0, k_op = 4, B' = 0 + 1 + 64/64
    = 2, so t_glb = max(5W + 4W, 0 + 1 + 2*5*W) = 10W + 1, write-back = 1,
    phase 1 = 30W; one stage of one block per CU (256 blocks on 256 CUs):
    4 launches x (40W + 2) cycles at 1 Hz, W = waves of one workgroup."""
    p = _p("jacobi-2d", (256, 256, 1), 4, 1, 64, 4, 1, 1)
    d = otlib.derive(p)
    assert d.grid_blocks[0] * d.grid_blocks[1] == 256
    W = math.ceil(d.threads_per_block / 64)
    a = _arch(num_cu=256.0, clock_hz=1.0, c_load=1.0, c_op=1.0, b_gmem=0.0, b_smem=0.0, l_gmem=0.0, l_smem=0.0,
              launch_penalty=0.0, l1_bytes=0.0, sfu_cycles=0.0)
    assert ottune.analytic_seconds(p, 32, a) == 4 * (40 * W + 2)


@pytest.mark.parametrize("name,n3,blk", [("jacobi-2d", (512, 384, 1), (32, 8, 1, 2)),
                                         ("jacobi-3d", (128, 96, 64), (32, 4, 16, 2)),
                                         ("rician-denoise-2d", (300, 200, 1), (16, 8, 1, 1))])
def test_analytic_phase_decomposition(otlib, name, n3, blk):
    """Stage time = phase 1 + global phase + (T-1) on-chip phases + write-back,
    so with no launch penalty and steps a multiple of T:
    full = (global-only) + (on-chip-only) - (prologue-only)."""
    a = _arch(launch_penalty=0.0)
    x, y, z, e = blk
    t = {}
    for pl in (0, 1, 2, 3):
        t[pl] = ottune.analytic_seconds(_p(name, n3, 12, 3, x, y, z, e, p=pl), 48, a)
    assert t[0] == pytest.approx(t[2] + t[3] - t[1], rel=1e-12)
    assert t[3] > t[1] and t[2] > t[1]


def test_analytic_scaling_in_steps_and_launch_penalty(otlib):
    a = ottune.default_arch()
    p8 = _p("jacobi-3d", (128, 128, 128), 8, 4, 32, 4, 16, 2)
    p16 = _p("jacobi-3d", (128, 128, 128), 16, 4, 32, 4, 16, 2)
    assert ottune.analytic_seconds(p16, 64, a) == pytest.approx(2 * ottune.analytic_seconds(p8, 64, a), rel=1e-12)
    a2 = _arch(launch_penalty=a.launch_penalty + 2.4e9 * 1e-6)
    # 2 launches each 1 us slower
    assert ottune.analytic_seconds(p8, 64, a2) - ottune.analytic_seconds(p8, 64, a) == pytest.approx(2e-6)
    # T=1: no on-chip phase, so the full kernel equals the global-only one
    q = _p("jacobi-3d", (128, 128, 128), 8, 1, 32, 4, 16, 2)
    q2 = _p("jacobi-3d", (128, 128, 128), 8, 1, 32, 4, 16, 2, p=2)
    assert ottune.analytic_seconds(q, 64, a) == pytest.approx(ottune.analytic_seconds(q2, 64, a), rel=1e-12)


def test_tune_entry_points_reject_bad_arguments(otlib):
    with pytest.raises(otlib.OtError):
        ottune.analytic_seconds(_p("jacobi-2d", (64, 64, 1), 4, 1, 4096, 4, 1, 1), 32)


def test_space_holds_all_2d_kernel_families(otlib):
    """2D spaces carry the LDS tiles, the wave-per-tile kernel (x tile + 2rT = 64,
    footprint rows 16/24/32/64) and the row-streaming kernel (chunks of >= 64 rows);
    Rician's streaming kernel stops at T = 5."""
    for name, R in (("jacobi-2d", 1), ("fdtd-2d", 1), ("rician-denoise-2d", 2)):
        sp = ottune.space(_p(name, (2048, 2048, 1), 24, 1, 16, 4, 1, 1))
        keys = [(q.time_tile, q.block_x, q.block_y, q.elems_per_thread) for q in sp]
        wave = [k for k in keys if k[1] + 2 * R * k[0] == 64]
        assert any(k[2] >= 64 for k in wave) and any(k[2] < 64 for k in wave), name
        assert any(k[1] + 2 * R * k[0] != 64 for k in keys), name
        for q in sp:
            d = otlib.derive(q)
            is_wave = q.block_x + 2 * R * q.time_tile == 64 and (q.block_y >= 64 or q.block_y + 2 * R * q.time_tile
                                                                in (16, 24, 32, 64))
            if is_wave and not (name == "rician-denoise-2d" and q.block_y >= 64 and q.time_tile > 5):
                assert d.lds_bytes == 0 and d.threads_per_block == 256, (name, q.time_tile, q.block_y)
        if name == "rician-denoise-2d":
            assert not [k for k in wave if k[2] >= 64 and k[0] > 5]


_FAKE_SWEEP = r'''
import csv, json, os, sys, time
rows = json.load(open(sys.argv[1]))
path, hang_at, crash_at = sys.argv[2], int(sys.argv[3]), int(sys.argv[4])
cols = ["Stencil", "Nx", "Ny", "Nz", "Time Steps", "Geometry", "Time Tile Size", "Block Size X",
        "Block Size Y", "Block Size Z", "Elements/Thread"]
done = set()
if os.path.exists(path):
    with open(path) as f:
        done = {tuple(r[c] for c in cols) for r in csv.DictReader(f)}
for i, r in enumerate(rows):
    if tuple(str(r[c]) for c in cols) in done:
        continue
    if i == hang_at:
        time.sleep(600)   # a kernel that never returns
    if i == crash_at:
        sys.exit(3)       # a process that dies on this configuration
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(",".join(r) + "\n")
        f.write(",".join(str(v) for v in r.values()) + "\n")
'''


def test_supervised_sweep_kills_a_hung_child_and_resumes(otlib, tmp_path):
    """The exhaustive sweep's watchdog (the reference's 15 s per-run watchdog,
    scripts/utils.py:130-145), exercised on the CPU with a stand-in for the
    sweep child: configuration 7 never returns and configuration 20 kills its
    process. Each is recorded once with Status hang / crash, a fresh child
    resumes, and every other configuration is measured exactly once."""
    p = _p("jacobi-1d", (4096, 1, 1), 8, 1, 64, 1, 1, 1)
    sp = ottune.space(p)
    rows = []
    for i, q in enumerate(sp):
        k = ottune.config_key(q)
        r = dict(zip(["Stencil", "Nx", "Ny", "Nz", "Time Steps", "Geometry", "Time Tile Size", "Block Size X",
                      "Block Size Y", "Block Size Z", "Elements/Thread"], k))
        full = {c: r.get(c, 0) for c in ottune.CSV_COLUMNS}
        full["EventElapsed"] = 1e-3 + 1e-6 * ((i * 37) % 101)
        full["Status"] = "ok"
        rows.append(full)
    jf = tmp_path / "rows.json"
    jf.write_text(__import__("json").dumps(rows))
    child = tmp_path / "fake_sweep.py"
    child.write_text(_FAKE_SWEEP)
    csv_path = str(tmp_path / "sweep.csv")
    import sys
    cmd = [sys.executable, str(child), str(jf), csv_path, "7", "20"]
    msgs = []
    best, best_s, info = ottune.exhaustive_supervised(p, csv_path, watchdog_s=0.5, startup_s=10.0, child_cmd=cmd,
                                                      log=msgs.append)
    assert info["restarts"] == 2 and len(msgs) == 2
    assert info["hung"] == [ottune.config_key(sp[7]), ottune.config_key(sp[20])]
    got = ottune.read_csv(csv_path)
    assert len(got) == len(sp)
    status = {ottune._row_key(r): r["Status"] for r in got}
    assert status[ottune.config_key(sp[7])] == "hang" and status[ottune.config_key(sp[20])] == "crash"
    assert sum(v == "ok" for v in status.values()) == len(sp) - 2
    ok = [(float(r["EventElapsed"]), i) for i, r in enumerate(rows) if i not in (7, 20)]
    assert best_s == pytest.approx(min(ok)[0])
    assert ottune.config_key(best) == ottune.config_key(sp[min(ok)[1]])
