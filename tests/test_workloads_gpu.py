"""Parity at the measured workload sizes, over all time steps (MI355X).

For every grid workload that fits one GPU (Jacobi-2D 1024^2 x 64, Jacobi-3D
256^3 x 32, Poisson-3D with its right-hand side 256^3 x 32, FDTD-2D 2048^2 x 100
with the Gaussian pulse, with the reference's Hz sign and with the curl's,
gradient-3D 256^3 x 64 on the noisy sinusoid, and cs-3d 256^3 x 64), the
engine runs the whole solve at the tile configuration the benchmark ships,
from the benchmark's own seeded inputs, and the result must
equal the FP32 oracle after all steps bit for bit (so within the 1e-5
max-relative bar). Each case also prints its error against the FP64 oracle.

The batched denoising workloads (256 images of 64..1024 x 64..768, 50
iterations, Rician and TV) must give, bit for bit, the same images batched as
launched one image at a time, and a sample of images must equal the FP32
oracle.
"""
import numpy as np
import pytest

import oracle
import otbatch
import otlib
import workloads
from helpers import MAX_REL

pytestmark = pytest.mark.gpu


def _fp64_error(st, rw, ro, steps, gpu, variant=0):
    ref = oracle.run(st, [x.astype(np.float64) for x in rw], [x.astype(np.float64) for x in ro], steps=steps,
                     dtype=np.float64, variant=variant)
    m = [oracle.compare(r, g) for r, g in zip(ref, gpu)]
    return max(x["max_rel"] for x in m), max(x["l2_rel"] for x in m)


CASES = [(name, 0) for name in workloads.WORKLOADS] + [("fdtd2d", otlib.VARIANT_FDTD_PHYSICAL)]


@pytest.mark.parametrize("name,variant", CASES, ids=[n + ("-physical" if v else "") for n, v in CASES])
def test_workload_all_steps_equal_the_oracle(otlib, name, variant, capsys):
    w = workloads.WORKLOADS[name]
    p = workloads.plan_params(w, w["tiled"])
    p.variant = variant
    d = otlib.derive(p)
    rw, ro = workloads.make_inputs(w, d)
    sl = otlib.interior_slices(d)
    print(f"workload {name}: {w['stencil']} {w['n']} x {w['steps']} steps at {w['tiled']}, variant {variant}",
          flush=True)
    res = otlib.run_host(p, rw, ro)
    gpu = [x[sl] for x in res.fields]
    st = p.stencil
    rw_i, ro_i = [x[sl].copy() for x in rw], [x[sl].copy() for x in ro]
    ref = oracle.run(st, [x.copy() for x in rw_i], ro_i, steps=w["steps"], variant=variant)
    for k, (r, g) in enumerate(zip(ref, gpu)):
        m = oracle.compare(r, g)
        assert m["same_nonfinite"] and m["max_rel"] <= MAX_REL, (k, m)
        assert m["bitwise_equal"], (k, m)
    err64, l2_64 = _fp64_error(st, rw_i, ro_i, w["steps"], gpu, variant)
    with capsys.disabled():  # shown in the run's output, also under -q
        print(f"\nworkload {name}{'-physical' if variant else ''}: bitwise equal to the FP32 oracle after "
              f"{w['steps']} steps; against the FP64 oracle max relative error {err64:.3e}, L2 relative "
              f"{l2_64:.3e}", flush=True)
    assert np.isfinite(err64)


def _image(arena, im):
    return otbatch.image_view(arena, im)[None].copy()


@pytest.mark.parametrize("name", list(workloads.DEFAULT_BATCH))
def test_batch_workload_equals_per_image_and_oracle(otlib, name, capsys):
    ims, total, widths, heights, rw, ro, coef = workloads.batch_inputs(name)
    st = otlib.RICIAN2D if name == "batch-rician" else otlib.TV2D
    T, bx, by, e = workloads.DEFAULT_BATCH[name]
    steps = 50
    p = otlib.make_params(st, steps=steps, T=T, x=bx, y=by, e=e, coef=coef)
    print(f"{name}: {len(ims)} images, {steps} iterations, T={T} {bx}x{by * e} e{e}", flush=True)
    batched, tm = otbatch.run_batch_host(p, ims, rw, ro, use_graph=True)
    per_image, _ = otbatch.run_batch_host(p, ims, rw, ro, per_image=True)
    assert tm.launches == -(-steps // T)
    for im in ims:
        assert np.array_equal(otbatch.image_view(batched[0], im).view(np.uint32),
                              otbatch.image_view(per_image[0], im).view(np.uint32)), (im.width, im.height)
    # the oracle on a sample: the largest, the smallest and six drawn images
    order = np.argsort([im.width * im.height for im in ims])
    rng = np.random.default_rng(5)
    sample = sorted({int(order[0]), int(order[-1])} | set(int(i) for i in rng.choice(len(ims), 6, replace=False)))
    worst64 = 0.0
    for i in sample:
        im = ims[i]
        rw_i, ro_i = [_image(x, im) for x in rw], [_image(x, im) for x in ro]
        ref = oracle.run(st, [x.copy() for x in rw_i], ro_i, steps=steps, coef=coef)[0]
        got = otbatch.image_view(batched[0], im)[None]
        m = oracle.compare(ref, got)
        assert m["bitwise_equal"], (i, im.width, im.height, m)
        ref64 = oracle.run(st, [x.astype(np.float64) for x in rw_i], [x.astype(np.float64) for x in ro_i],
                           steps=steps, coef=coef, dtype=np.float64)[0]
        worst64 = max(worst64, oracle.compare(ref64, got)["max_rel"])
    with capsys.disabled():
        print(f"\n{name}: batched == per-image on all {len(ims)} images; images {sample} bitwise equal to the "
              f"FP32 oracle; max relative error against the FP64 oracle {worst64:.3e}", flush=True)
