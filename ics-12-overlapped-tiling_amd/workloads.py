"""The measured workloads: grids, time steps, seeded inputs and the tile
configurations the exhaustive autotuner picked for each on an MI355X.

Shared by the benchmark and the workload-size parity tests, so both run the
very same inputs through the very same plans.
"""
from __future__ import annotations

import numpy as np

import otlib

# Grid workloads: stencil, interior extents (nx, ny, nz), time steps, the best
# tile configuration the exhaustive autotuner (ot_tune_exhaustive) found for
# exactly this workload on an MI355X (T, block x, block y rows, block z,
# elements per thread), the best T=1 configuration of the same sweep, and the
# algorithmic HBM bytes per point per launch (every rw field read and written
# once, every ro field read once).
WORKLOADS = {
    "jacobi3d": dict(stencil="jacobi-3d", n=(256, 256, 256), steps=32,
                     tiled=dict(T=3, x=58, y=6, z=19, e=1), t1=dict(T=1, x=62, y=6, z=8, e=1),
                     bytes_pp=8, data="uniform [0,1) interior from glibc srand(123456), Dirichlet-0 boundary"),
    "jacobi2d": dict(stencil="jacobi-2d", n=(1024, 1024, 1), steps=64,
                     tiled=dict(T=8, x=48, y=8, z=1, e=1), t1=dict(T=1, x=16, y=16, z=1, e=4),
                     bytes_pp=8, data="uniform [0,1) interior from glibc srand(123456), Dirichlet-0 boundary"),
    "fdtd2d": dict(stencil="fdtd-2d", n=(2048, 2048, 1), steps=100,
                   tiled=dict(T=7, x=50, y=10, z=1, e=1), t1=dict(T=1, x=62, y=14, z=1, e=1),
                   bytes_pp=24, data="Ex = Ey = 0, Hz Gaussian pulse (amplitude 1, sigma 8) at the centre"),
    "gradient3d": dict(stencil="gradient-3d", n=(256, 256, 256), steps=64,
                       tiled=dict(T=2, x=60, y=8, z=14, e=1), t1=dict(T=1, x=62, y=6, z=14, e=1),
                       bytes_pp=8, data="sin(2 pi x/256) sin(2 pi y/256) sin(2 pi z/256) + U(-1e-3, 1e-3), seed 7"),
    "poisson3d-rhs": dict(stencil="poisson-3d-rhs", n=(256, 256, 256), steps=32,
                          tiled=dict(T=2, x=60, y=8, z=22, e=1), t1=dict(T=1, x=62, y=6, z=64, e=1),
                          bytes_pp=12, data="u0 = 0, f uniform [-1, 1) from seed 11"),
    # compute-bound (4 rsqrt and 5 divisions per point): the sweep's best tile is
    # a T=1 tile, so the time-tiled and T=1 lines coincide
    "cs3d": dict(stencil="cs-3d", n=(256, 256, 256), steps=64,
                 tiled=dict(T=1, x=62, y=6, z=19, e=1), t1=dict(T=1, x=62, y=6, z=19, e=1),
                 bytes_pp=8, data="uniform [0,1) interior from glibc srand(123456), Dirichlet-0 boundary"),
}

# Batch workloads (256 images of 64..1024 x 64..768, 50 iterations): (T, block x,
# block y rows, elements per thread) of the batched work-queue launch, and of the
# per-image baseline (one launch per image per time tile) at the configuration
# that is fastest for per-image launches (both from sweeps of the batch on an
# MI355X).
DEFAULT_BATCH = {"batch-rician": (2, 64, 8, 4), "batch-tv": (3, 58, 26, 1)}
PER_IMAGE_BEST = {"batch-rician": (4, 32, 8, 2), "batch-tv": (5, 54, 6, 1)}
BATCH_SEED = 77
# kernel scalars: Rician SIGMA2, LAMBDA, DT as the reference's literals
# (rician-denoise-2d.cpp:153-160); TV with lambda (its alpha) 0.1
BATCH_COEF = {"batch-rician": [0.0025, 0.065, 5.0, 0.0], "batch-tv": [0.1, 0.0, 0.0, 0.0]}


def plan_params(w, cfg, n3=None, steps=None):
    """ot_params of grid workload w at tile configuration cfg."""
    n3 = n3 or w["n"]
    st = otlib.stencil_id(w["stencil"])
    return otlib.make_params(st, nx=n3[0], ny=n3[1], nz=n3[2], problem_size=n3[0], steps=steps or w["steps"],
                             T=cfg["T"], x=cfg["x"], y=cfg["y"], z=cfg["z"], e=cfg["e"])


def make_inputs(w, d):
    """Host arrays (layout of derived d) for the rw and ro fields of workload w."""
    st = otlib.stencil_id(w["stencil"])
    if st in (otlib.J3D, otlib.J2D, otlib.CS3D):
        return [otlib.fill_reference_uniform(d, 123456)], []
    if st == otlib.FDTD2D:
        z = otlib.empty_field(d)
        return [z, z.copy(), otlib.fill_gaussian_pulse(d, 1.0, 8.0)], []
    if st == otlib.G3D:
        return [otlib.fill_sinusoid_noise(d, 7, 1e-3)], []
    if st == otlib.P3D7:
        return [otlib.empty_field(d)], [otlib.fill_uniform(d, 11, -1.0, 1.0)]
    raise ValueError(w["stencil"])


def batch_inputs(name, n_images=256):
    """(images, arena elements, widths, heights, rw arenas, ro arenas, coef) of a
    batch workload: rectangle phantoms with Rician noise sigma 0.05; u0 = f;
    TV with V = 1 and F = f."""
    import otbatch

    widths, heights = otbatch.draw_sizes(n_images, seed=BATCH_SEED)
    ims, total = otbatch.layout(widths, heights)
    f = otbatch.phantom_arena(ims, total, seed=BATCH_SEED, sigma=0.05)
    if name == "batch-rician":
        ro = [f]
    else:
        ro = [np.ones(total, np.float32), f]
    return ims, total, widths, heights, [f.copy()], ro, list(BATCH_COEF[name])
