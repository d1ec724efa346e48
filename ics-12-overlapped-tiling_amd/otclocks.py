#!/usr/bin/env python3
"""Per-wave timeline of the wave-pipelined jacobi-3d kernel from its clock
records (OT_VARIANT_DUMP_CLOCKS, read with ot_plan_clocks).

Runs the jacobi-3d workload (256^3 by default) at a wave tile for a few tenths
of a second to settle the clocks, then one solve of 10 full time tiles with
clock records, and summarises the last launch: when the waves start and end
(s_memrealtime, 100 MHz, one time base for the whole device), how long each
spends in the pipeline fill, the steady-state blocks and the rest (s_memtime
cycles, converted at the measured shader clock), how many waves each SIMD held,
how the lifetime of a wave depends on its start order on its SIMD (the
hardware issues the oldest wave first unless the kernel balances progress), how
the dies (XCDs) compare, and which z-chunks hold the waves that end last.

  python3 ics-12-overlapped-tiling_amd/otclocks.py                 # T=3 58x6x19
  python3 ics-12-overlapped-tiling_amd/otclocks.py --no-balance    # oldest-first issue
  python3 ics-12-overlapped-tiling_amd/otclocks.py --offset-mb 19  # fields 19 MiB into their allocations
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import otlib  # noqa: E402
import workloads  # noqa: E402


class ClockRecord(C.Structure):
    _fields_ = [("block_x", C.c_int32), ("block_y", C.c_int32), ("wave", C.c_int32), ("hw_id", C.c_uint32),
                ("xcc_id", C.c_uint32), ("valid", C.c_uint32), ("clock", C.c_uint64 * 5),
                ("realtime", C.c_uint64 * 2)]


class _At:
    """A field at a byte offset inside a larger device allocation."""

    def __init__(self, ptr: int):
        self.ptr = ptr


def records(n: int, T: int, bx: int, by: int, bz: int, variant: int, settle_s: float, offset_bytes: int = 0):
    """Clock records of the last launch; offset_bytes places both fields that
    far into allocations of field size + offset (to see whether timing follows
    the addresses)."""
    w = workloads.WORKLOADS["jacobi3d"]
    cfg = dict(T=T, x=bx, y=by, z=bz, e=1)
    p = workloads.plan_params(w, cfg, (n, n, n), 10 * T)
    p.variant = variant | otlib.VARIANT_DUMP_CLOCKS
    plan = otlib.Plan(p)
    d = plan.derived
    rw, _ = workloads.make_inputs(w, d)
    allocs = []
    A, B = [], []
    for x in rw:
        x = np.ascontiguousarray(x)
        for side in (A, B):
            buf = otlib.DeviceBuffer(x.nbytes + offset_bytes)
            allocs.append(buf)
            side.append(_At(buf.ptr + offset_bytes))
            otlib.check(otlib.lib().ot_memcpy_h2d(side[-1].ptr, x.ctypes.data, x.nbytes))
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < settle_s:
        for _ in range(8):
            plan.run(A, B, [], enqueue_only=True)
        otlib.check(otlib.lib().ot_dev_sync())
    plan.run(A, B, [])
    cnt = C.c_int32()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, None, 0, C.byref(cnt)))
    recs = (ClockRecord * cnt.value)()
    otlib.check(otlib.lib().ot_plan_clocks(plan.handle, recs, cnt.value, C.byref(cnt)))
    plan.close()
    for b in allocs:
        b.free()
    return [r for r in recs if r.valid]


def pct(v, unit=1.0):
    return " / ".join(f"{np.percentile(v, q) * unit:.2f}" for q in (0, 10, 50, 90, 100))


def summarise(recs) -> str:
    rt = np.array([[r.realtime[0], r.realtime[1]] for r in recs], dtype=np.int64)
    ck = np.array([list(r.clock) for r in recs], dtype=np.int64)
    hw = np.array([r.hw_id for r in recs], dtype=np.int64)
    xcc = np.array([r.xcc_id & 0xF for r in recs], dtype=np.int64)
    start = (rt[:, 0] - rt[:, 0].min()) / 100.0  # us
    end = (rt[:, 1] - rt[:, 0].min()) / 100.0
    life = end - start
    mhz = np.median((ck[:, 4] - ck[:, 0]) / np.maximum(life, 1e-3))  # cycles per us
    fill, steady, drain = [(ck[:, i + 1] - ck[:, i]) / mhz for i in range(3)]
    # the wave's SIMD: XCC, shader engine, shader array, CU, SIMD
    simd = (((xcc * 8 + ((hw >> 13) & 7)) * 2 + ((hw >> 12) & 1)) * 16 + ((hw >> 8) & 0xF)) * 4 + ((hw >> 4) & 3)
    keys, inv, per = np.unique(simd, return_inverse=True, return_counts=True)
    simd_end = np.zeros(len(keys))
    np.maximum.at(simd_end, inv, end)
    order = np.zeros(len(recs), dtype=np.int64)  # start order of the wave on its SIMD
    for k in range(len(keys)):
        idx = np.where(inv == k)[0]
        order[idx[np.argsort(start[idx], kind="stable")]] = np.arange(len(idx))
    lines = [
        f"waves {len(recs)} on {len(keys)} SIMDs (waves per SIMD: "
        + ", ".join(f"{c} x{n}" for c, n in zip(*np.unique(per, return_counts=True))) + ")",
        f"launch span {end.max():.2f} us (first wave start to last wave end); shader clock {mhz:.0f} MHz",
        "percentiles p0 / p10 / p50 / p90 / p100, us:",
        f"  wave start          {pct(start)}",
        f"  wave end            {pct(end)}",
        f"  wave lifetime       {pct(life)}",
        f"  pipeline fill (P0-P1) {pct(fill)}",
        f"  steady blocks (P1-P2) {pct(steady)}",
        f"  last planes (P2-P3)   {pct(drain)}",
        f"  SIMD end            {pct(simd_end)}",
        "mean lifetime by start order on the SIMD: "
        + ", ".join(f"{o}: {life[order == o].mean():.2f} us" for o in range(order.max() + 1)),
    ]
    lines.append("by XCD (XCC_ID): mean lifetime / last end, us; shader clock, MHz: "
                 + ", ".join(f"{x}: {life[xcc == x].mean():.2f} / {end[xcc == x].max():.2f}; "
                             f"{np.median((ck[xcc == x, 4] - ck[xcc == x, 0]) / np.maximum(life[xcc == x], 1e-3)):.0f}"
                             for x in np.unique(xcc)))
    # by z-chunk (the record's wave field): which chunks hold the waves that end last
    bz = np.array([r.wave for r in recs], dtype=np.int64)
    last = np.argsort(end, kind="stable")[-max(1, len(recs) // 20):]
    lines.append("by z-chunk: mean lifetime / last end, us; waves among the last-ending 5 %:")
    for z in np.unique(bz):
        m = bz == z
        lines.append(f"  chunk {z:3d}: {m.sum():5d} waves  {life[m].mean():6.2f} / {end[m].max():6.2f}"
                     f"  {int(np.sum(bz[last] == z))}")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--T", type=int, default=3)
    ap.add_argument("--bx", type=int, default=58)
    ap.add_argument("--by", type=int, default=6)
    ap.add_argument("--bz", type=int, default=19)
    ap.add_argument("--no-balance", action="store_true", help="oldest-first issue (OT_VARIANT_NO_BALANCE)")
    ap.add_argument("--settle", type=float, default=0.3)
    ap.add_argument("--offset-mb", type=int, default=0,
                    help="place both fields this many MiB into larger allocations")
    a = ap.parse_args()
    variant = otlib.VARIANT_NO_BALANCE if a.no_balance else 0
    recs = records(a.n, a.T, a.bx, a.by, a.bz, variant, a.settle, a.offset_mb << 20)
    print(f"jacobi-3d {a.n}^3, T={a.T} tile {a.bx}x{a.by}x{a.bz}, "
          f"{'oldest-first issue' if a.no_balance else 'default issue policy'}; last launch of a "
          f"{10 * a.T}-step solve")
    print(summarise(recs), flush=True)


if __name__ == "__main__":
    main()
