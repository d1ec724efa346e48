"""Batches of independent 2D images through the engine's C ABI (ot_batch_*).

A batch packs images of uneven size into one arena per field; every
time-tile launch is one persistent kernel draining a device work queue of
(image, tile) items, and the whole loop can be replayed from a hipGraph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import otlib
from otlib import DeviceBuffer, Image, Params, Timing, _ptr_array, check, lib

RUN_PER_IMAGE = 4


def layout(widths, heights):
    """(images ctypes array, arena elements) for the given extents."""
    n = len(widths)
    w = (C.c_int32 * n)(*map(int, widths))
    h = (C.c_int32 * n)(*map(int, heights))
    ims = (Image * n)()
    total = C.c_int64()
    check(lib().ot_batch_layout(n, w, h, ims, C.byref(total)))
    return ims, int(total.value)


def phantom_arena(images, total: int, seed: int, sigma: float = 0.05) -> np.ndarray:
    a = np.zeros(total, np.float32)
    check(lib().ot_fill_rician_phantom(a.ctypes.data_as(C.POINTER(C.c_float)), len(images), images, seed, sigma))
    return a


def image_view(arena: np.ndarray, im: Image) -> np.ndarray:
    """(height, width) view of one image inside an arena."""
    return arena[im.offset: im.offset + im.pitch * im.height].reshape(im.height, im.pitch)[:, :im.width]


def draw_sizes(n: int, seed: int, wmin=64, wmax=1024, hmin=64, hmax=768):
    rng = np.random.default_rng(seed)
    return rng.integers(wmin, wmax + 1, n), rng.integers(hmin, hmax + 1, n)


class Batch:
    def __init__(self, params: Params, images):
        self.params = params
        self.images = images
        h = C.c_void_p()
        check(lib().ot_batch_create(C.byref(params), len(images), images, C.byref(h)))
        self.handle = h
        items, wgs, launches = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().ot_batch_info(h, C.byref(items), C.byref(wgs), C.byref(launches)))
        self.items, self.workgroups, self.launches = items.value, wgs.value, launches.value

    def run(self, rw_a, rw_b, ro=(), use_graph=False, per_image=False, stream=None) -> Timing:
        """per_image=True: the baseline of one per-grid launch per image and
        time tile instead of the work-queue kernel."""
        a = _ptr_array([b.ptr for b in rw_a])
        b = _ptr_array([x.ptr for x in rw_b])
        r = _ptr_array([x.ptr for x in ro]) if ro else None
        tm = Timing()
        mode = (otlib.RUN_GRAPH if use_graph else 0) | (RUN_PER_IMAGE if per_image else 0)
        check(lib().ot_batch_run(self.handle, a, b, r, stream, mode, C.byref(tm)))
        return tm

    def close(self):
        if self.handle:
            lib().ot_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch_host(params: Params, images, rw_arenas, ro_arenas=(), use_graph=False, per_image=False):
    """Upload arenas, run the batch, download the rw arenas."""
    bt = Batch(params, images)
    a = [DeviceBuffer.from_array(x) for x in rw_arenas]
    b = [DeviceBuffer.from_array(x) for x in rw_arenas]
    ro = [DeviceBuffer.from_array(x) for x in ro_arenas]
    tm = bt.run(a, b, ro, use_graph=use_graph, per_image=per_image)
    src = a if tm.final_buffer == 0 else b
    out = [buf.download(x.shape) for buf, x in zip(src, rw_arenas)]
    for buf in a + b + ro:
        buf.free()
    bt.close()
    return out, tm


def sweep(stencil: str, n_images: int = 256, iterations: int = 50, seed: int = 77):
    """Time every candidate tile of a denoising batch both ways: the batched
    work-queue launch (hipGraph replay, best of 3) and one per-grid launch per
    image and time tile (faster of eager and graph). Candidates are LDS tiles
    (T 1..6, 9 block shapes, 1/2/4 elements per thread) and the tiles that fill
    a wavefront (x tile + 2rT = 64, footprint rows 16, 24 or 32). Returns one dict
    per configuration that builds."""
    import itertools
    import sys
    import time

    st = otlib.stencil_id(stencil)
    R = otlib.stencil_info(st).radius
    w, h = draw_sizes(n_images, seed=seed)
    ims, total = layout(w, h)
    f = phantom_arena(ims, total, seed=seed, sigma=0.05)
    coef = [0.0025, 0.065, 5.0, 0.0] if st == otlib.RICIAN2D else [0.1, 0.0, 0.0, 0.0]
    ro = [f] if st == otlib.RICIAN2D else [np.ones(total, np.float32), f]
    A, B = DeviceBuffer.from_array(f), DeviceBuffer.from_array(f)
    RO = [DeviceBuffer.from_array(x) for x in ro]

    def best_ms(bt, graph, per_image, reps):
        bt.run([A], [B], RO, use_graph=graph, per_image=per_image)
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            bt.run([A], [B], RO, use_graph=graph, per_image=per_image)
            best = min(best, time.perf_counter() - t0)
        return best * 1e3

    cfgs = [(T, bx, by, e) for T, (bx, by), e in itertools.product(
        range(1, 7), [(32, 8), (32, 16), (64, 8), (64, 16), (32, 32), (48, 32), (64, 32), (128, 8), (16, 16)],
        (1, 2, 4)) if by % e == 0]
    for T in range(1, 9):
        x = 64 - 2 * R * T
        for fr in (16, 24, 32):
            if x > 0 and fr > 2 * R * T:
                cfgs.append((T, x, fr - 2 * R * T, 1))
    out = []
    try:
        for T, bx, rows, e in cfgs:
            try:
                bt = Batch(otlib.make_params(st, steps=iterations, T=T, x=bx, y=rows // e, e=e, coef=coef), ims)
            except otlib.OtError:
                continue
            try:
                fr = rows + 2 * R * T  # footprint rows of a wave tile
                wave = bx + 2 * R * T == 64 and e == 1 and fr in (16, 24, 32, 64) and 3 * fr <= 128
                out.append(dict(T=T, tile=[bx, rows], e=e, wave=int(wave),
                                batched_ms=round(best_ms(bt, True, False, 3), 4),
                                per_image_ms=round(min(best_ms(bt, False, True, 1), best_ms(bt, True, True, 1)), 4)))
                print(f"{name} T={T} {bx}x{rows} e{e}: batched {out[-1]['batched_ms']} ms, "
                      f"per image {out[-1]['per_image_ms']} ms", file=sys.stderr, flush=True)
            finally:
                bt.close()
    finally:
        for buf in [A, B] + RO:
            buf.free()
    return out


if __name__ == "__main__":
    import json
    import sys

    for name in (sys.argv[1:] or ["rician-denoise-2d", "cdsc-tv-update-2d"]):
        res = sweep(name)
        for key in ("batched_ms", "per_image_ms"):
            top = sorted(res, key=lambda r: r[key])[:5]
            print(json.dumps({"stencil": name, "configs": len(res), "fastest_by": key, "top": top}), flush=True)
