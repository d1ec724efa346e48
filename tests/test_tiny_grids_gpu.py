"""Degenerate and tiny grids: domains of one to a few points per axis, far
smaller than the tile, through every kernel family (1D LDS, 2D LDS,
wave-per-tile and row-streaming, 3D LDS and wave-pipelined), against the FP32
oracle bit for bit. The whole domain then sits inside one tile's halo: every
lane, row and plane outside it must read the Dirichlet 0 and store nothing.

The CPU test pins the oracle on a one-point domain, where a Jacobi update with
a zero boundary reduces to the centre term: v <- c * v in FP32, one rounding
per step (the six or so zero neighbours add exactly nothing).
"""
import numpy as np
import pytest

import oracle
import otlib
from helpers import check_against_oracle, domain_shape, random_inputs

S = otlib
# (stencil, (nx, ny, nz), steps, T, block, elements per thread)
TINY = [
    # 1D LDS
    (S.J1D3, (1, 1, 1), 5, 1, (64, 1, 1), 1),
    (S.J1D3, (2, 1, 1), 7, 4, (64, 1, 1), 2),
    (S.J1D5, (3, 1, 1), 6, 3, (64, 1, 1), 1),
    (S.J1D7, (5, 1, 1), 9, 2, (32, 1, 1), 3),
    # 2D LDS
    (S.J2D, (1, 1, 1), 5, 2, (16, 16, 1), 1),
    (S.J2D, (2, 3, 1), 7, 3, (16, 8, 1), 2),
    (S.J2D, (5, 1, 1), 6, 4, (8, 8, 1), 1),
    (S.G2D, (1, 2, 1), 5, 2, (16, 16, 1), 1),
    (S.RICIAN2D, (1, 3, 1), 5, 1, (16, 16, 1), 1),
    (S.FDTD2D, (2, 3, 1), 6, 3, (16, 8, 1), 3),
    # 2D wave-per-tile
    (S.J2D, (1, 1, 1), 5, 2, (60, 12, 1), 1),
    (S.J2D, (3, 2, 1), 7, 4, (56, 8, 1), 1),
    (S.P2D, (2, 2, 1), 5, 3, (58, 6, 1), 1),
    (S.G2D, (3, 1, 1), 5, 2, (60, 8, 1), 1),
    (S.TV2D, (2, 3, 1), 5, 2, (60, 4, 1), 1),
    (S.RICIAN2D, (2, 2, 1), 5, 2, (56, 8, 1), 1),
    (S.FDTD2D, (1, 1, 1), 5, 1, (62, 30, 1), 1),
    # 2D row-streaming
    (S.J2D, (1, 5, 1), 6, 1, (62, 64, 1), 1),
    (S.J2D, (4, 3, 1), 9, 8, (48, 64, 1), 1),
    (S.TV2D, (3, 3, 1), 5, 3, (58, 64, 1), 1),
    (S.RICIAN2D, (3, 2, 1), 6, 3, (52, 64, 1), 1),
    (S.FDTD2D, (3, 2, 1), 7, 3, (58, 64, 1), 1),
    # 3D LDS
    (S.J3D, (1, 1, 1), 5, 2, (16, 8, 8), 1),
    (S.J3D, (3, 2, 1), 7, 3, (16, 8, 8), 2),
    (S.G3D, (2, 3, 4), 5, 2, (16, 8, 8), 1),
    (S.P3D27, (1, 2, 1), 4, 1, (8, 8, 4), 1),
    # 3D wave-pipelined, at the bench tiles and others
    (S.J3D, (1, 1, 1), 5, 3, (58, 6, 19), 1),
    (S.J3D, (5, 4, 3), 7, 3, (58, 6, 19), 1),
    (S.J3D, (2, 1, 6), 8, 4, (56, 4, 16), 1),
    (S.G3D, (2, 3, 4), 5, 2, (60, 8, 14), 1),
    (S.P3D7, (2, 3, 4), 5, 2, (60, 8, 22), 1),
    (S.P3D27, (2, 3, 4), 5, 2, (60, 8, 8), 1),
    (S.CS3D, (2, 3, 4), 5, 2, (60, 8, 8), 1),
    (S.CS3D, (1, 1, 1), 4, 3, (58, 6, 16), 1),
]


def _id(c):
    st, n3, steps, T, block, e = c
    return f"{S.STENCILS[st]}-n{'x'.join(map(str, n3))}-t{steps}-T{T}-b{'x'.join(map(str, block))}-e{e}"


def _params(case):
    st, (nx, ny, nz), steps, T, block, e = case
    return otlib.make_params(st, nx=nx, ny=ny, nz=nz, problem_size=nx, steps=steps, T=T, x=block[0],
                             y=block[1], z=block[2], e=e, geometry="tile")


@pytest.mark.gpu
@pytest.mark.parametrize("case", TINY, ids=[_id(c) for c in TINY])
def test_tiny_grid_matches_oracle(otlib, case):
    p = _params(case)
    d = otlib.derive(p)
    assert list(d.grid_blocks) == [1, 1, 1]  # the domain fits one tile
    rw, ro = random_inputs(case[0], domain_shape(d), seed=case[0] * 10 + case[3])
    print(f"case {_id(case)}", flush=True)
    check_against_oracle(p, rw, ro)


@pytest.mark.parametrize("st,coef", [(S.J1D3, 0.333), (S.J1D5, 0.2), (S.J1D7, 0.2), (S.J2D, 0.2), (S.J3D, 0.143)])
def test_oracle_one_point_domain_is_centre_scaling(st, coef):
    """jacobi-1d.cpp:631-639, jacobi-1d-5pt.cpp:505-515, jacobi-1d-7pt.cpp:513-523,
    jacobi-2d.cpp:824-828, jacobi-3d.cpp:729-746: the update is a coefficient
    times the sum of the taps; with a one-point domain and a zero boundary the
    sum is the centre value itself."""
    v0 = np.float32(0.7377)
    got = oracle.run(st, [np.full((1, 1, 1), v0, np.float32)], [], steps=9, coef=[0, 0, 0, 0], variant=0)[0]
    v = v0
    for _ in range(9):
        v = np.float32(np.float32(coef) * v)
    assert got.reshape(-1)[0] == v
