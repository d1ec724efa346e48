"""Timing probes through the C ABI (MI355X): the reference's --time-only modes
of the 2D LDS kernel (jacobi-2d.cpp:258-266, :296-341, :368-433, :462-481) and
its per-warp phase clocks (--dump-clocks, :202-205, :453-509), the latter also
for the wave-pipelined jacobi-3d kernel.

  * clock stamps change no result bit, one record per wavefront of the last
    launch, every stamp sequence in phase order;
  * gmem copies the loaded footprint's tile interior back out (so a whole
    solve leaves the input unchanged), instr and smem store nothing;
  * the 3D wave kernel's issue-priority balancing changes no result bit.
"""
import ctypes as C

import numpy as np
import pytest

import otlib
from helpers import domain_shape, random_inputs, run_gpu

pytestmark = pytest.mark.gpu


class ClockRecord(C.Structure):
    _fields_ = [("block_x", C.c_int32), ("block_y", C.c_int32), ("wave", C.c_int32), ("hw_id", C.c_uint32),
                ("xcc_id", C.c_uint32), ("valid", C.c_uint32), ("clock", C.c_uint64 * 5),
                ("realtime", C.c_uint64 * 2)]


CASES = [(otlib.J2D, (300, 211, 1), 10, 4, (32, 8, 1), 2), (otlib.G2D, (257, 190, 1), 9, 3, (16, 16, 1), 1),
         (otlib.J2D, (128, 96, 1), 7, 1, (64, 4, 1), 4)]


def _p(st, n3, steps, T, blk, ept, variant=0):
    return otlib.make_params(st, nx=n3[0], ny=n3[1], nz=1, problem_size=n3[0], steps=steps, T=T, x=blk[0],
                             y=blk[1], z=1, e=ept, variant=variant)


@pytest.mark.parametrize("case", CASES, ids=[f"{otlib.STENCILS[c[0]]}-T{c[3]}-e{c[5]}" for c in CASES])
def test_clock_stamps_leave_results_unchanged(otlib, case):
    st, n3, steps, T, blk, ept = case
    p0 = _p(st, n3, steps, T, blk, ept)
    pc = _p(st, n3, steps, T, blk, ept, variant=otlib.VARIANT_DUMP_CLOCKS)
    d = otlib.derive(p0)
    assert d.lds_bytes > 0
    rw, ro = random_inputs(st, domain_shape(d), seed=3)
    print(f"probe clocks {otlib.STENCILS[st]} {n3} T={T} {blk} e{ept}", flush=True)
    ref, _ = run_gpu(p0, rw, ro)
    plan = otlib.Plan(pc)
    full = otlib.empty_field(plan.derived)
    full[otlib.interior_slices(plan.derived)] = rw[0]
    A, B = otlib.DeviceBuffer.from_array(full), otlib.DeviceBuffer.from_array(full)
    tm = plan.run([A], [B])
    out = (B if tm.final_buffer else A).download(full.shape)[otlib.interior_slices(plan.derived)]
    assert np.array_equal(out.view(np.uint32), ref[0].view(np.uint32))
    n = C.c_int32()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, None, 0, C.byref(n)))
    blocks = d.grid_blocks[0] * d.grid_blocks[1]
    assert n.value == blocks * 16
    recs = (ClockRecord * n.value)()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, recs, n.value, C.byref(n)))
    # the last launch advances steps - (launches - 1) T steps: its workgroup
    # covers that launch's footprint
    t_last = steps - (d.launches - 1) * T
    fx, fy = blk[0] + 2 * t_last, blk[1] * ept + 2 * t_last
    waves = -(-(fx * -(-fy // ept)) // 64)
    seen = set()
    for i, r in enumerate(recs):
        if not r.valid:
            assert i % 16 >= waves
            continue
        assert r.wave == i % 16 < waves
        c = list(r.clock)
        assert c[0] <= c[1] <= c[2] <= c[3] <= c[4], c
        assert r.realtime[1] >= r.realtime[0] > 0
        seen.add((r.block_x, r.block_y, r.wave))
    assert len(seen) == blocks * waves
    A.free()
    B.free()
    plan.close()


@pytest.mark.parametrize("case", CASES, ids=[f"{otlib.STENCILS[c[0]]}-T{c[3]}-e{c[5]}" for c in CASES])
@pytest.mark.parametrize("mode,bit", [("gmem", otlib.VARIANT_TIME_ONLY_GMEM), ("instr", otlib.VARIANT_TIME_ONLY_INSTR),
                                      ("smem", otlib.VARIANT_TIME_ONLY_SMEM)])
def test_time_only_modes(otlib, case, mode, bit):
    st, n3, steps, T, blk, ept = case
    p = _p(st, n3, steps, T, blk, ept, variant=bit)
    d = otlib.derive(p)
    rw, _ = random_inputs(st, domain_shape(d), seed=4)
    print(f"time-only {mode} {otlib.STENCILS[st]} {n3} T={T} {blk} e{ept}", flush=True)
    plan = otlib.Plan(p)
    full = otlib.empty_field(d)
    full[otlib.interior_slices(d)] = rw[0]
    marker = np.full(full.shape, -7.0, np.float32)
    A, B = otlib.DeviceBuffer.from_array(full), otlib.DeviceBuffer.from_array(marker)
    tm = plan.run([A], [B])
    assert tm.event_s > 0
    a = A.download(full.shape)
    b = B.download(full.shape)
    if mode == "gmem":
        # every launch copies its input's interior to its output
        assert np.array_equal((B if tm.final_buffer else A).download(full.shape)[otlib.interior_slices(d)], rw[0])
    else:
        assert np.array_equal(a, full) and np.array_equal(b, marker)
    A.free()
    B.free()
    plan.close()


def test_probes_are_refused_where_the_reference_has_none(otlib):
    for st, n3, blk in [(otlib.P2D, (64, 64, 1), (16, 16, 1)), (otlib.J2D, (256, 64, 1), (60, 12, 1))]:
        with pytest.raises(otlib.OtError) as e:
            otlib.Plan(_p(st, n3, 4, 2, blk, 1, variant=otlib.VARIANT_TIME_ONLY_GMEM))
        assert e.value.status == otlib.OT_EUNSUPPORTED
    # 3D: clock records for jacobi-3d wave tiles only, no --time-only
    for st, blk, bit in [(otlib.P3D7, (60, 4, 8), otlib.VARIANT_DUMP_CLOCKS),
                         (otlib.J3D, (16, 8, 8), otlib.VARIANT_DUMP_CLOCKS),
                         (otlib.J3D, (60, 4, 8), otlib.VARIANT_TIME_ONLY_INSTR)]:
        p = otlib.make_params(st, nx=64, ny=64, nz=64, problem_size=64, steps=4, T=2, x=blk[0], y=blk[1], z=blk[2],
                              e=1, variant=bit)
        with pytest.raises(otlib.OtError) as e:
            otlib.Plan(p)
        assert e.value.status == otlib.OT_EUNSUPPORTED


def _j3d(n3, steps, T, blk, variant=0):
    return otlib.make_params(otlib.J3D, nx=n3[0], ny=n3[1], nz=n3[2], problem_size=n3[0], steps=steps, T=T,
                             x=blk[0], y=blk[1], z=blk[2], e=1, variant=variant)


def _solve(p, rw):
    plan = otlib.Plan(p)
    d = plan.derived
    full = otlib.empty_field(d)
    full[otlib.interior_slices(d)] = rw[0]
    A, B = otlib.DeviceBuffer.from_array(full), otlib.DeviceBuffer.from_array(full)
    tm = plan.run([A], [B])
    out = (B if tm.final_buffer else A).download(full.shape)[otlib.interior_slices(d)]
    A.free()
    B.free()
    return plan, out


# wave tiles (x + 2T = 64): the last launch of 7 steps at T=3 is a T=1 launch
CASES_3D = [((100, 70, 45), 7, 3, (58, 6, 10)), ((130, 64, 64), 6, 2, (60, 4, 19)), ((64, 50, 33), 5, 4, (56, 8, 9))]


@pytest.mark.parametrize("case", CASES_3D, ids=[f"T{c[2]}-{c[0][0]}x{c[0][1]}x{c[0][2]}" for c in CASES_3D])
def test_3d_wave_clock_records(otlib, case):
    n3, steps, T, blk = case
    p0 = _j3d(n3, steps, T, blk)
    d = otlib.derive(p0)
    rw, _ = random_inputs(otlib.J3D, domain_shape(d), seed=6)
    print(f"3D clocks {n3} T={T} {blk} steps {steps}", flush=True)
    _, ref = _solve(p0, rw)
    plan, out = _solve(_j3d(n3, steps, T, blk, variant=otlib.VARIANT_DUMP_CLOCKS), rw)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    g = plan.derived.grid_blocks
    n = C.c_int32()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, None, 0, C.byref(n)))
    assert n.value == g[0] * g[1] * g[2]
    recs = (ClockRecord * n.value)()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, recs, n.value, C.byref(n)))
    seen = set()
    for i, r in enumerate(recs):
        assert r.valid == 1
        # record i is wave i: x tile fastest, then y tile, then z-chunk
        assert (r.block_x, r.block_y, r.wave) == (i % g[0], (i // g[0]) % g[1], i // (g[0] * g[1]))
        c = list(r.clock)
        assert c[0] <= c[1] <= c[2] <= c[3] <= c[4], c
        assert r.realtime[1] >= r.realtime[0] > 0
        seen.add((r.block_x, r.block_y, r.wave))
    assert len(seen) == n.value
    plan.close()


def test_3d_wave_balancing_changes_no_result_bit(otlib):
    n3, steps, T, blk = (256, 96, 80), 8, 3, (58, 6, 19)
    p = _j3d(n3, steps, T, blk)
    d = otlib.derive(p)
    rw, _ = random_inputs(otlib.J3D, domain_shape(d), seed=8)
    print(f"3D balance {n3} T={T} {blk}", flush=True)
    a = _solve(p, rw)
    b = _solve(_j3d(n3, steps, T, blk, variant=otlib.VARIANT_NO_BALANCE), rw)
    a[0].close()
    b[0].close()
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
