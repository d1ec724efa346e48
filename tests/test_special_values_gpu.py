"""IEEE special values through every kernel family, against the FP32 oracle,
with two input sets: "wide", fields sprinkled with denormals (down to the
smallest, 2^-149), signed zeros and huge finite values (1e18, whose squares in
the gradient stencils reach 1e36 and whose Rician correction overflows to inf
and then NaN); and "tiny", the same fields scaled by 1e-38 (most values
denormal, so the stencils' arithmetic runs in the denormal range) and
sprinkled with zeros and the smallest denormals; and "nonfinite", sprinkled
with infinities, NaN and values near the largest float, whose sums overflow.
Points outside the domain must stay +0 whatever a masked update computes
there (infinite, NaN or negative sums included), as on the reference's ring.
The engine
must give the same bits at every point where the oracle's value is not NaN,
and NaN exactly where the oracle's is (NaN payloads differ between the CPU,
whose invalid operations give a negative quiet NaN, and the GPU).
"""
import numpy as np
import pytest

import oracle
import otlib
from helpers import domain_shape, random_inputs, run_gpu

pytestmark = pytest.mark.gpu

S = otlib
# name: (scale of the field, values sprinkled in, fraction of points replaced)
SETS = {
    "wide": (1.0, np.array([0.0, -0.0, 1e-40, -1e-40, 2.0 ** -149, -(2.0 ** -149), 3e-39, 1e18, -1e18], np.float32), 0.03),
    "tiny": (1e-38, np.array([0.0, -0.0, 2.0 ** -149, -(2.0 ** -149), 2.0 ** -140], np.float32), 0.03),
    "nonfinite": (1.0, np.array([np.inf, -np.inf, np.nan, -3e38, 3e38], np.float32), 0.001),
}
# (stencil, (nx, ny, nz), steps, T, block, elements per thread): one case per kernel family and stencil
CASES = [
    (S.J1D3, (1000, 1, 1), 13, 4, (64, 1, 1), 3),
    (S.J1D7, (777, 1, 1), 11, 5, (32, 1, 1), 4),
    (S.J2D, (100, 77, 1), 9, 4, (16, 16, 1), 1),            # 2D LDS
    (S.G2D, (90, 61, 1), 7, 2, (16, 16, 1), 1),
    (S.J2D, (130, 77, 1), 9, 2, (60, 12, 1), 1),            # 2D wave-per-tile
    (S.P2D, (90, 61, 1), 7, 3, (58, 26, 1), 1),
    (S.G2D, (90, 61, 1), 9, 4, (56, 8, 1), 1),
    (S.TV2D, (100, 70, 1), 11, 5, (54, 6, 1), 1),
    (S.RICIAN2D, (90, 61, 1), 7, 2, (56, 8, 1), 1),
    (S.FDTD2D, (130, 70, 1), 15, 7, (50, 10, 1), 1),
    (S.J2D, (200, 150, 1), 9, 2, (60, 64, 1), 1),           # 2D row-streaming
    (S.FDTD2D, (130, 170, 1), 8, 3, (58, 64, 1), 1),
    (S.J3D, (40, 37, 35), 7, 2, (16, 8, 8), 2),             # 3D LDS
    (S.P3D27, (36, 30, 29), 5, 2, (16, 8, 8), 1),
    (S.J3D, (70, 20, 60), 7, 3, (58, 6, 16), 1),            # 3D wave-pipelined (steady blocks too)
    (S.G3D, (64, 20, 50), 5, 2, (60, 8, 12), 1),
    (S.P3D7, (70, 25, 29), 9, 4, (56, 4, 8), 1),
    (S.P3D27, (70, 25, 29), 5, 2, (60, 8, 8), 1),
    (S.CS3D, (70, 30, 29), 5, 2, (60, 8, 8), 1),
]


def _id(c):
    st, n3, steps, T, block, e = c
    return f"{S.STENCILS[st]}-n{'x'.join(map(str, n3))}-T{T}-b{'x'.join(map(str, block))}-e{e}"


def _sprinkle(a: np.ndarray, rng, scale, specials, frac) -> np.ndarray:
    a = (a * np.float32(scale)).astype(np.float32)
    flat = a.reshape(-1)
    idx = rng.choice(flat.size, size=max(1, int(frac * flat.size)), replace=False)
    flat[idx] = rng.choice(specials, size=idx.size)
    return a


def _same(ref: np.ndarray, got: np.ndarray):
    rb = np.asarray(ref, np.float32).view(np.uint32)
    gb = np.asarray(got, np.float32).view(np.uint32)
    rn, gn = np.isnan(ref), np.isnan(got)
    return bool(np.array_equal(rn, gn) and np.array_equal(rb[~rn], gb[~gn])), int(rn.sum())


@pytest.mark.parametrize("which", list(SETS))
@pytest.mark.parametrize("case", CASES, ids=[_id(c) for c in CASES])
def test_special_values_match_oracle(otlib, case, which):
    st, (nx, ny, nz), steps, T, block, e = case
    p = otlib.make_params(st, nx=nx, ny=ny, nz=nz, problem_size=nx, steps=steps, T=T, x=block[0], y=block[1],
                          z=block[2], e=e, geometry="tile")
    d = otlib.derive(p)
    rng = np.random.default_rng(1000 + st * 10 + T)
    rw, ro = random_inputs(st, domain_shape(d), seed=st * 100 + T)
    scale, specials, frac = SETS[which]
    rw = [_sprinkle(x, rng, scale, specials, frac) for x in rw]
    print(f"case {_id(case)} {which} steps={steps}", flush=True)
    gpu, _ = run_gpu(p, rw, ro)
    ref = oracle.run(st, [x.copy() for x in rw], [x for x in ro], steps=steps, coef=list(p.coef), variant=p.variant)
    for k, (g, r) in enumerate(zip(gpu, ref)):
        same, nans = _same(r, g)
        assert same, (k, nans, int(np.sum(np.asarray(r, np.float32).view(np.uint32) != np.asarray(g, np.float32).view(np.uint32))))
