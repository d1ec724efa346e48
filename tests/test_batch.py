"""Batched uneven 2D images: one persistent work-queue launch per time tile.

Invariant: the batched result of every image equals, bit for bit, the same
image run alone (per-grid launches) and the FP32 oracle on that image."""
import numpy as np
import pytest

import oracle
import otbatch
import otlib


def test_layout_packs_aligned_arena(otlib):
    ims, total = otbatch.layout([5, 64, 17], [3, 2, 9])
    assert [im.pitch for im in ims] == [16, 64, 32]
    assert all(im.offset % 16 == 0 for im in ims)
    assert ims[1].offset >= ims[0].offset + 16 * 3 and total >= ims[2].offset + 32 * 9


def test_phantom_generator_deterministic(otlib):
    ims, total = otbatch.layout([40, 23], [30, 11])
    a = otbatch.phantom_arena(ims, total, seed=9)
    b = otbatch.phantom_arena(ims, total, seed=9)
    c = otbatch.phantom_arena(ims, total, seed=10)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for im in ims:
        v = otbatch.image_view(a, im)
        assert np.all(np.isfinite(v)) and np.all(v >= 0)
        pad = a[im.offset: im.offset + im.pitch * im.height].reshape(im.height, im.pitch)[:, im.width:]
        assert np.all(pad == 0)


def _batch_case(st, T, bx, by, e, steps, seed, coef=(0, 0, 0, 0)):
    rng = np.random.default_rng(seed)
    widths = rng.integers(20, 141, 9)
    heights = rng.integers(15, 101, 9)
    widths[0], heights[0] = 3, 2  # a tiny image: a single partial tile
    ims, total = otbatch.layout(widths, heights)
    f = otbatch.phantom_arena(ims, total, seed)
    nrw = 3 if st == otlib.FDTD2D else 1
    if st == otlib.FDTD2D:
        rw = [rng.random(total).astype(np.float32) - 0.5 for _ in range(3)]
        ro = []
    elif st == otlib.TV2D:
        rw = [f.copy()]
        ro = [np.ones(total, np.float32) + 0.5 * rng.random(total).astype(np.float32), f.copy()]
    elif st == otlib.RICIAN2D:
        rw, ro = [f.copy()], [f.copy()]
    else:
        rw, ro = [f.copy()], []
    p = otlib.make_params(st, steps=steps, T=T, x=bx, y=by, e=e, coef=list(coef))
    out, tm = otbatch.run_batch_host(p, ims, rw, ro)
    out_g, _ = otbatch.run_batch_host(p, ims, rw, ro, use_graph=True)
    out_pi, _ = otbatch.run_batch_host(p, ims, rw, ro, per_image=True)
    for a, b, c in zip(out, out_g, out_pi):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        for im in ims:  # padding columns are not written by the per-image launches either
            assert np.array_equal(otbatch.image_view(a, im).view(np.uint32), otbatch.image_view(c, im).view(np.uint32))
    for im in ims:
        rw_i = [otbatch.image_view(x, im)[None].copy() for x in rw]
        ro_i = [otbatch.image_view(x, im)[None].copy() for x in ro]
        ref = oracle.run(st, [x.copy() for x in rw_i], ro_i, steps=steps, coef=list(coef))
        # the same image alone through the per-grid kernels
        pi = otlib.make_params(st, nx=im.width, ny=im.height, n=im.width, steps=steps, T=T, x=bx, y=by, e=e,
                               coef=list(coef))
        d = otlib.derive(pi)
        alone = otlib.run_host(pi, [_embed(d, x) for x in rw_i], [_embed(d, x) for x in ro_i])
        for k in range(nrw):
            got = otbatch.image_view(out[k], im)[None]
            m = oracle.compare(ref[k], got)
            assert m["bitwise_equal"], (im.width, im.height, m)
            assert np.array_equal(got.view(np.uint32),
                                  alone.fields[k][otlib.interior_slices(d)].view(np.uint32))
    return tm


def _embed(d, interior):
    a = otlib.empty_field(d)
    a[otlib.interior_slices(d)] = interior
    return a


@pytest.mark.gpu
@pytest.mark.parametrize("st,T,bx,by,e", [
    (otlib.RICIAN2D, 3, 16, 8, 2), (otlib.RICIAN2D, 1, 32, 16, 1), (otlib.TV2D, 4, 32, 4, 4),
    (otlib.J2D, 5, 16, 16, 1), (otlib.FDTD2D, 2, 16, 8, 2), (otlib.G2D, 3, 32, 8, 1), (otlib.P2D, 2, 16, 8, 2),
    # tiles that fill a wavefront (x tile + 2rT = 64): the wave-per-tile batch kernel
    (otlib.RICIAN2D, 2, 56, 24, 1), (otlib.TV2D, 3, 58, 26, 1), (otlib.J2D, 4, 56, 8, 1),
    (otlib.FDTD2D, 2, 60, 12, 1), (otlib.G2D, 5, 54, 22, 1), (otlib.P2D, 8, 48, 48, 1),
])
def test_batch_equals_per_image_and_oracle(otlib, st, T, bx, by, e):
    coef = (0.1, 0, 0, 0) if st == otlib.TV2D else (0, 0, 0, 0)
    tm = _batch_case(st, T, bx, by, e, steps=11, seed=st * 10 + T, coef=coef)
    assert tm.launches == -(-11 // T)
