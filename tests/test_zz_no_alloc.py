"""The allocation-free batch entry point (ot_batch_run_static) through the C
ABI (MI355X): its work queue lives in the library's fixed per-GPU scratch, and
its results must equal ot_batch_create + ot_batch_run's bit for bit, call after
call, also when a call's queue is smaller than the one before it (stale items
and counters of the previous call stay in the scratch)."""
import ctypes as C

import numpy as np
import pytest

import otbatch
import otlib
from otlib import DeviceBuffer, Timing, _ptr_array, check, lib

pytestmark = pytest.mark.gpu


def _inputs(st, n_images, wmax, hmax, seed):
    rng = np.random.default_rng(seed)
    widths, heights = rng.integers(20, wmax + 1, n_images), rng.integers(15, hmax + 1, n_images)
    ims, total = otbatch.layout(widths, heights)
    f = otbatch.phantom_arena(ims, total, seed)
    ro = [f.copy()] if st == otlib.RICIAN2D else [np.ones(total, np.float32) + 0.25, f.copy()]
    return ims, total, [f.copy()], ro


def _static(p, ims, rw, ro):
    A = [DeviceBuffer.from_array(x) for x in rw]
    B = [DeviceBuffer.from_array(x) for x in rw]
    R = [DeviceBuffer.from_array(x) for x in ro]
    tm = Timing()
    check(lib().ot_batch_run_static(C.byref(p), len(ims), ims, _ptr_array([b.ptr for b in A]),
                                    _ptr_array([b.ptr for b in B]), _ptr_array([b.ptr for b in R]), None,
                                    C.byref(tm)))
    check(lib().ot_dev_sync())
    out = [buf.download(x.shape) for buf, x in zip(B if tm.final_buffer else A, rw)]
    for buf in A + B + R:
        buf.free()
    return out, tm


CALLS = [  # (stencil, images, max width, max height, steps, T, block x, block y rows, ept, seed)
    (otlib.RICIAN2D, 96, 700, 500, 13, 2, 64, 8, 4, 1),   # the Rician batch tile, LDS kernel
    (otlib.RICIAN2D, 11, 160, 120, 9, 2, 64, 8, 4, 2),    # smaller queue, same stencil and tile
    (otlib.TV2D, 40, 400, 300, 11, 3, 58, 26, 1, 3),      # the TV batch tile, wave kernel
    (otlib.TV2D, 5, 90, 70, 7, 3, 58, 26, 1, 4),          # smaller again
]


def test_static_batch_equals_batch_run_call_after_call(otlib):
    for st, n, wmax, hmax, steps, T, bx, by, e, seed in CALLS:
        ims, total, rw, ro = _inputs(st, n, wmax, hmax, seed)
        coef = [0.1, 0, 0, 0] if st == otlib.TV2D else [0, 0, 0, 0]
        p = otlib.make_params(st, steps=steps, T=T, x=bx, y=by, e=e, coef=coef)
        print(f"static batch {otlib.STENCILS[st]}: {n} images, {steps} steps, T={T} {bx}x{by * e} e{e}", flush=True)
        ref, tref = otbatch.run_batch_host(p, ims, rw, ro)
        got, tm = _static(p, ims, rw, ro)
        assert tm.launches == tref.launches and tm.final_buffer == tref.final_buffer
        for a, b in zip(ref, got):
            for im in ims:
                assert np.array_equal(otbatch.image_view(a, im).view(np.uint32),
                                      otbatch.image_view(b, im).view(np.uint32)), (im.width, im.height)


def test_static_batch_refuses_queues_beyond_its_scratch(otlib):
    st = otlib.J2D
    ims, total = otbatch.layout([8] * 5000, [8] * 5000)
    z = [DeviceBuffer(4 * total), DeviceBuffer(4 * total)]
    p = otlib.make_params(st, steps=4, T=2, x=16, y=8, e=1)
    tm = Timing()
    status = lib().ot_batch_run_static(C.byref(p), len(ims), ims, _ptr_array([z[0].ptr]), _ptr_array([z[1].ptr]),
                                       None, None, C.byref(tm))
    assert status == otlib.OT_ERESOURCE and b"use ot_batch_create" in lib().ot_last_error()
    # 2000 images of 512 x 512 in 16 x 8 tiles: 4.1M items, beyond 8 MiB
    ims, total = otbatch.layout([512] * 2000, [512] * 2000)
    status = lib().ot_batch_run_static(C.byref(p), len(ims), ims, _ptr_array([z[0].ptr]), _ptr_array([z[1].ptr]),
                                       None, None, C.byref(tm))
    assert status == otlib.OT_ERESOURCE
    for b in z:
        b.free()
