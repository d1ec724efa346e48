#!/usr/bin/env python3
"""Benchmark of the overlapped time-tiled stencil engine on one MI355X.

Workload (default): Jacobi-3D 7-point FP32 on a 256^3 grid, 32 time steps,
interior values uniform [0,1) from a seed, Dirichlet-0 boundary, time-tiled with
the tile configuration below. One bench "step" is one whole 32-time-step solve
of the grid (ceil(32/T) kernel launches replayed from a hipGraph), with the
fields resident in HBM.

Prints one JSON line with the metric (GStencil/s = grid points x time steps /
second), the roofline of the dominant kernel (algorithmic HBM bytes per launch
= 8 B x points: each launch reads and writes every point once, divided by the
average launch duration measured with HIP events on the launch stream), the
same workload at T=1 for the time-tiling gain, and the CPU baseline (the
oracle's restatement of the reference CPU loop, single-threaded, on the same
workload).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ics-12-overlapped-tiling_amd"))

import numpy as np  # noqa: E402

import otlib  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Best tile configurations for Jacobi-3D 256^3 found by the engine's exhaustive
# sweep (T, block x, block y rows, block z, elements per thread) and T=1.
DEFAULT_TILED = dict(T=2, x=32, y=4, z=32, e=4)
DEFAULT_T1 = dict(T=1, x=64, y=4, z=32, e=4)
# batch workloads: (T, block x, block y rows, elements per thread)
DEFAULT_BATCH = {"batch-rician": (3, 32, 8, 4), "batch-tv": (2, 64, 4, 4)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "jacobi3d_256_traffic.json")


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed solves")
    ap.add_argument("--warmup", type=int, default=3, help="untimed solves")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--time-steps", type=int, default=32)
    ap.add_argument("--T", type=int, default=DEFAULT_TILED["T"])
    ap.add_argument("--bx", type=int, default=DEFAULT_TILED["x"])
    ap.add_argument("--by", type=int, default=DEFAULT_TILED["y"], help="block y in rows of ept")
    ap.add_argument("--bz", type=int, default=DEFAULT_TILED["z"])
    ap.add_argument("--ept", type=int, default=DEFAULT_TILED["e"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-t1", action="store_true", help="skip the T=1 comparison")
    ap.add_argument("--workload", default="jacobi3d",
                    choices=["jacobi3d", "batch-rician", "batch-tv"],
                    help="jacobi3d: the headline line; batch-*: 256 uneven images, batched persistent "
                         "launches vs one launch per image and time tile")
    ap.add_argument("--images", type=int, default=256)
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes) and exit")
    return ap.parse_args()


def calibrate_copy(mib):
    nbytes = mib << 20
    a = otlib.DeviceBuffer(nbytes)
    b = otlib.DeviceBuffer(nbytes)
    otlib.check(otlib.lib().ot_dev_memset(a.ptr, 0, nbytes))
    for _ in range(5):
        otlib.check(otlib.lib().ot_memcpy_d2d(b.ptr, a.ptr, nbytes))
    otlib.check(otlib.lib().ot_dev_sync())
    print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 5}), flush=True)


def solve_rate(cfg, n, steps_t, K, W, init):
    """Time K whole solves (graph replay) after W warm-ups. Returns
    (seconds per solve, average launch seconds from the launch-stream events,
    launches per solve, eager per-launch average)."""
    p = otlib.make_params(otlib.J3D, n=n, steps=steps_t, T=cfg["T"], x=cfg["x"], y=cfg["y"], z=cfg["z"],
                          e=cfg["e"])
    plan = otlib.Plan(p)
    a = otlib.DeviceBuffer.from_array(init)
    b = otlib.DeviceBuffer.from_array(init)
    for _ in range(W):
        plan.run([a], [b], use_graph=True)
    otlib.check(otlib.lib().ot_dev_sync())
    ev_total = 0.0
    t0 = time.perf_counter()
    for _ in range(K):
        tm = plan.run([a], [b], use_graph=True)
        ev_total += tm.event_s
    otlib.check(otlib.lib().ot_dev_sync())
    t1 = time.perf_counter()
    launches = plan.derived.launches
    # eager pass with one event pair per launch (cross-check of the launch time)
    eager = []
    for _ in range(max(2, min(K, 5))):
        tm = plan.run([a], [b], time_launches=True)
        eager.append(tm.kernel_s / tm.launches)
    plan.close()
    a.free()
    b.free()
    return (t1 - t0) / K, ev_total / (K * launches), launches, float(np.median(eager))


def cpu_baseline(n, steps_t):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    rng = np.random.default_rng(12345)
    u = rng.random((n, n, n), dtype=np.float32)
    _, dt = oracle.run(otlib.J3D, [u], steps=steps_t, timed=True)
    return {
        "value": round(n ** 3 * steps_t / dt / 1e9, 4),
        "unit": "GStencil/s",
        "cores": 1,
        "kind": "port",
        "sample": f"whole workload: Jacobi-3D {n}^3, {steps_t} time steps, oracle C restatement of the "
                  f"reference CPU loop (jacobi-3d.cpp:729-746), gcc -O2 -ffp-contract=off, 1 thread of "
                  f"{os.cpu_count()} ({platform.processor() or platform.machine()}), {dt:.2f} s",
    }


def bench_batch(args, stencil_name):
    """Batched Rician / TV denoising of uneven images: one persistent
    work-queue launch per time tile (hipGraph replay) against the same images
    advanced by one per-grid launch per image and time tile."""
    import otbatch

    st = otlib.RICIAN2D if stencil_name == "rician" else otlib.TV2D
    widths, heights = otbatch.draw_sizes(args.images, seed=77)
    ims, total = otbatch.layout(widths, heights)
    f = otbatch.phantom_arena(ims, total, seed=77, sigma=0.05)
    coef = [0.05, 0.065, 5.0, 0.0] if st == otlib.RICIAN2D else [0.1, 0.0, 0.0, 0.0]
    T, bx, by, e = (args.T, args.bx, args.by, args.ept)
    p = otlib.make_params(st, steps=args.iterations, T=T, x=bx, y=by, e=e, coef=coef)
    ro_host = [f] if st == otlib.RICIAN2D else [np.ones(total, np.float32), f]
    A = otlib.DeviceBuffer.from_array(f)
    B = otlib.DeviceBuffer.from_array(f)
    RO = [otlib.DeviceBuffer.from_array(x) for x in ro_host]
    bt = otbatch.Batch(p, ims)

    def timed(use_graph, per_image, reps):
        for _ in range(args.warmup):
            bt.run([A], [B], RO, use_graph=use_graph, per_image=per_image)
        t0 = time.perf_counter()
        for _ in range(reps):
            bt.run([A], [B], RO, use_graph=use_graph, per_image=per_image)
        otlib.check(otlib.lib().ot_dev_sync())
        return (time.perf_counter() - t0) / reps

    batched = timed(True, False, args.steps)
    batched_eager = timed(False, False, args.steps)
    per_image = timed(False, True, max(2, args.steps // 2))
    per_image_graph = timed(True, True, max(2, args.steps // 2))
    pixels = float(sum(int(im.width) * int(im.height) for im in ims))
    out = {
        "metric": f"batched {stencil_name} denoise Gpixel-iterations/s",
        "value": round(pixels * args.iterations / batched / 1e9, 3),
        "unit": "Gpixel-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(batched * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: rectangle phantoms + Rician noise sigma 0.05, seed 77",
        "config": {"workload": f"{args.images} images {int(widths.min())}..{int(widths.max())} x "
                               f"{int(heights.min())}..{int(heights.max())}, {args.iterations} iterations",
                   "pixels": int(pixels), "time_tile": T, "tile": [bx, by * e], "elements_per_thread": e,
                   "items": bt.items, "workgroups": bt.workgroups, "launches_per_step": bt.launches},
        "batched_eager_ms": round(batched_eager * 1e3, 4),
        "per_image_ms": round(per_image * 1e3, 4),
        "per_image_graph_ms": round(per_image_graph * 1e3, 4),
        "per_image_launches": int(len(ims) * bt.launches),
        "speedup_vs_per_image": round(per_image / batched, 3),
        "speedup_vs_per_image_graph": round(per_image_graph / batched, 3),
    }
    bt.close()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload.startswith("batch-"):
        if args.T == DEFAULT_TILED["T"] and args.bx == DEFAULT_TILED["x"]:
            args.T, args.bx, args.by, args.ept = DEFAULT_BATCH[args.workload]
        bench_batch(args, args.workload.split("-", 1)[1])
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this engine's path is single-GPU; "
              f"see DESIGN.md for how the stencil would shard over several GPUs.", file=sys.stderr)
        sys.exit(2)
    n, steps_t = args.n, args.time_steps
    tiled = dict(T=args.T, x=args.bx, y=args.by, z=args.bz, e=args.ept)
    d0 = otlib.derive(otlib.make_params(otlib.J3D, n=n, steps=steps_t, T=1, x=32, y=8, z=32, e=1))
    init = otlib.fill_uniform(d0, 2024, 0.0, 1.0)

    sec, launch_s, launches, eager_launch_s = solve_rate(tiled, n, steps_t, args.steps, args.warmup, init)
    points = float(n) ** 3
    gst = points * steps_t / sec / 1e9
    bytes_per_launch = 8.0 * points  # read + write every point once per launch
    achieved = bytes_per_launch / launch_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE))
        if tr.get("config") == {"n": n, **tiled}:
            traffic = tr.get("hbm_bytes_per_launch")
    out = {
        "metric": "GStencil/s & %HBM roofline; learned tile pick vs exhaustive autotune",
        "value": round(gst, 2),
        "unit": "GStencil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: uniform [0,1) interior from seed 2024, Dirichlet-0 boundary",
        "config": {"workload": f"jacobi-3d 7-point {n}^3, {steps_t} time steps, time-tiled",
                   "time_tile": tiled["T"], "tile": [tiled["x"], tiled["y"] * tiled["e"], tiled["z"]],
                   "elements_per_thread": tiled["e"], "launches_per_step": launches, "graph": True},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "ot::k3::k3d<J3D>", "avg_launch_us": round(launch_s * 1e6, 2),
                     "eager_avg_launch_us": round(eager_launch_s * 1e6, 2),
                     "bytes_per_launch": bytes_per_launch},
    }
    if not args.no_t1:
        sec1, launch1, l1, _ = solve_rate(DEFAULT_T1, n, steps_t, max(3, args.steps // 2), args.warmup, init)
        g1 = points * steps_t / sec1 / 1e9
        out["t1"] = {"value": round(g1, 2), "ms_per_step": round(sec1 * 1e3, 4),
                     "achieved_GBs": round(8.0 * points / launch1 / 1e9, 1),
                     "tile": [DEFAULT_T1["x"], DEFAULT_T1["y"] * DEFAULT_T1["e"], DEFAULT_T1["z"]]}
        out["speedup_vs_t1"] = round(gst / g1, 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, steps_t)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
