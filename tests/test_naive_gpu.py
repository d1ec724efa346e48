"""The non-tiled one-step-per-launch kernels (the drivers' --verify path and the
no-tiling baseline) against the CPU oracle, bit for bit."""
import pytest

import oracle
import otlib
from helpers import domain_shape, random_inputs, to_engine, from_engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("st", range(len(otlib.STENCILS)), ids=otlib.STENCILS)
def test_naive_matches_oracle(otlib, st):
    dims = otlib.stencil_info(st).dims
    n3 = {1: (700, 1, 1), 2: (61, 47, 1), 3: (23, 21, 19)}[dims]
    # cs-3d plans only on wave-pipelined tiles (x = 64 - 2T); the naive path
    # takes the plan's layout only
    bx, by = (62, 6) if st == otlib.CS3D else (16, 8)
    p = otlib.make_params(st, nx=n3[0], ny=n3[1], nz=n3[2], n=n3[0], steps=7, T=1, x=bx, y=by, z=8)
    d = otlib.derive(p)
    rw, ro = random_inputs(st, domain_shape(d), seed=40 + st)
    res = otlib.run_naive_host(p, [to_engine(d, x) for x in rw], [to_engine(d, x) for x in ro])
    ref = oracle.run(st, [x.copy() for x in rw], ro, steps=7)
    for g, r in zip(res.fields, ref):
        m = oracle.compare(r, from_engine(d, g))
        assert m["bitwise_equal"], m
