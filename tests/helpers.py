"""Shared helpers for the parity tests: layout conversion and one GPU-vs-oracle
comparison. The oracle (oracle/oracle.py) is the checker; the engine runs
through its C ABI (otlib)."""
import numpy as np

import oracle
import otlib

# Max relative error bar of FP32 results against the FP32 oracle after all
# time steps (FMA contraction would be the only allowed difference; the engine
# is built with contraction off, so results are expected to be bit-identical).
MAX_REL = 1e-5


def to_engine(d, interior: np.ndarray) -> np.ndarray:
    a = otlib.empty_field(d)
    a[otlib.interior_slices(d)] = interior
    return a


def from_engine(d, arr: np.ndarray) -> np.ndarray:
    return arr[otlib.interior_slices(d)]


def domain_shape(d):
    return (d.real_size[2], d.real_size[1], d.real_size[0])


def random_inputs(stencil: int, shape, seed: int = 1):
    """Seeded interiors (rw list, ro list) with the value ranges each stencil
    is exercised on."""
    rng = np.random.default_rng(seed)
    u = rng.random(shape, dtype=np.float32)
    if stencil == otlib.FDTD2D:
        return [rng.random(shape, dtype=np.float32) - 0.5 for _ in range(3)], []
    if stencil == otlib.TV2D:
        v = (0.5 + rng.random(shape, dtype=np.float32)).astype(np.float32)
        f = rng.random(shape, dtype=np.float32)
        return [u], [v, f]
    if stencil == otlib.RICIAN2D:
        f = rng.random(shape, dtype=np.float32)
        return [f.copy()], [f]
    if stencil == otlib.P3D7:
        f = (2 * rng.random(shape, dtype=np.float32) - 1).astype(np.float32)
        return [u], [f]
    return [u], []


def run_gpu(params, rw_int, ro_int=(), use_graph=False):
    d = otlib.derive(params)
    rw = [to_engine(d, x) for x in rw_int]
    ro = [to_engine(d, x) for x in ro_int]
    res = otlib.run_host(params, rw, ro, use_graph=use_graph)
    return [from_engine(d, x) for x in res.fields], res


def check_against_oracle(params, rw_int, ro_int=(), use_graph=False, require_bitwise=True):
    """Run the engine and the FP32 oracle on the same inputs; assert the bar.
    Returns (metrics per field, gpu fields, oracle fields)."""
    gpu, _ = run_gpu(params, rw_int, ro_int, use_graph=use_graph)
    ref = oracle.run(params.stencil, [x.copy() for x in rw_int], [x for x in ro_int],
                     steps=params.time_steps, coef=list(params.coef), variant=params.variant)
    metrics = []
    for g, r in zip(gpu, ref):
        m = oracle.compare(r, g)
        metrics.append(m)
        assert m["same_nonfinite"], m
        assert m["max_rel"] <= MAX_REL, m
        if require_bitwise:
            assert m["bitwise_equal"], m
    return metrics, gpu, ref
