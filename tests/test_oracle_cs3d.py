"""Pins of the CPU oracle's cs-3d update (runs on the CPU, no GPU needed).

cs-3d has no reference CPU loop (cs-3d.cpp:700-716 is a copy of Jacobi-3D), so
the oracle restates the kernel text (cs-3d.cpp:286-308). It is pinned by what
the formula fixes:
  * the tap set: perturbing one voxel of a random field changes the output at
    exactly the 13 points whose stencil holds that voxel (a wrong or swapped
    offset moves a point in or out of that set);
  * zero and constant fields (the update is a weighted mean of the centre and
    its six face neighbours, whose weights sum to the denominator);
  * every output lies between the least and greatest of those seven values
    for a positive field (weights >= 0);
  * hand-derived closed forms around a single spike on a constant field, where
    each of c1..c4 reduces to a / or b / sqrt(eps + k d^2) / 1.1 with k = 0, 1, 3
    (pins alpha, sumAref, eps, the factor 3 and which coefficient takes which
    differences).
"""
import math

import numpy as np
import pytest

import oracle

CS3D = 13
# (dz, dy, dx) of every value one output point reads (cs-3d.cpp:250-283 and
# :325-379, named val_{dz}_{dy}_{dx})
TAPS = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
        (-1, 1, 0), (-1, 0, 1), (1, -1, 0), (0, -1, 1), (1, 0, -1), (0, 1, -1)]
EPS, SUM_AREF, ALPHA = 1e-9, 1.1, 1.1


def test_cs3d_taps_are_the_thirteen_named_offsets():
    rng = np.random.default_rng(5)
    u = (0.5 + rng.random((9, 9, 9))).astype(np.float32)
    p = (4, 4, 4)
    v = u.copy()
    v[p] += np.float32(0.3)
    a = oracle.run(CS3D, [u], steps=1)[0]
    b = oracle.run(CS3D, [v], steps=1)[0]
    changed = {tuple(int(i) for i in idx) for idx in np.argwhere(a != b)}
    expect = {(p[0] - t[0], p[1] - t[1], p[2] - t[2]) for t in TAPS}
    assert changed == expect


def test_cs3d_zero_and_constant_fields():
    z = oracle.run(CS3D, [np.zeros((6, 7, 8), np.float32)], steps=3)[0]
    assert np.all(z == 0.0)
    c = np.float32(0.625)
    out = oracle.run(CS3D, [np.full((8, 9, 10), c, np.float32)], steps=1)[0]
    core = out[1:-1, 1:-1, 1:-1]   # points whose 13 taps all lie in the domain
    assert np.allclose(core, c, rtol=1e-6, atol=0)


def test_cs3d_output_is_a_weighted_mean_of_centre_and_face_neighbours():
    rng = np.random.default_rng(9)
    u = (0.1 + 0.9 * rng.random((10, 11, 12))).astype(np.float32)
    out = oracle.run(CS3D, [u], steps=1)[0].astype(np.float64)
    pu = np.pad(u.astype(np.float64), 1)   # Dirichlet 0 ring
    nz, ny, nx = u.shape
    faces = [pu[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
             for dz, dy, dx in TAPS[:7]]
    lo = np.minimum.reduce(faces)
    hi = np.maximum.reduce(faces)
    assert np.all(out >= lo * (1 - 1e-6)) and np.all(out <= hi * (1 + 1e-6))
    # and a point's diagonal taps matter: the output differs from the same
    # weighted mean with all weights equal
    assert not np.allclose(out, np.mean(faces, axis=0), rtol=1e-3)


def _spike(a, b):
    u = np.full((7, 7, 7), b, np.float32)
    u[3, 3, 3] = a
    return u, oracle.run(CS3D, [u], steps=1)[0].astype(np.float64)


@pytest.mark.parametrize("a,b", [(0.9, 0.5), (0.2, 0.7)])
def test_cs3d_spike_closed_forms(a, b):
    a32, b32 = float(np.float32(a)), float(np.float32(b))
    d = (a32 - b32) ** 2
    _, out = _spike(a, b)

    def mean(u0, w, vals):
        # (alpha u0 + 3 c1 (forward faces) + c2 z- + c3 y- + c4 x-) / (alpha + 3 c1 + c2 + c3 + c4)
        c1, c2, c3, c4 = w
        fz, fy, fx, zm, ym, xm = vals
        return (ALPHA * u0 + c1 * (fz + fy + fx) + c2 * zm + c3 * ym + c4 * xm) / \
            (ALPHA + 3 * c1 + c2 + c3 + c4)

    def c(u0, k):
        return u0 / math.sqrt(EPS + k * d) / SUM_AREF

    # at the spike: c1 sees three differences, c2..c4 one each
    exp = mean(a32, (c(a32, 3), c(a32, 1), c(a32, 1), c(a32, 1)), (b32,) * 6)
    assert out[3, 3, 3] == pytest.approx(exp, rel=2e-6)
    # spike = z-1 of (4,3,3): c2 sees three differences
    exp = mean(b32, (c(b32, 0), c(b32, 3), c(b32, 0), c(b32, 0)), (b32, b32, b32, a32, b32, b32))
    assert out[4, 3, 3] == pytest.approx(exp, rel=2e-6)
    # spike = y-1 of (3,4,3): c3 sees three differences
    exp = mean(b32, (c(b32, 0), c(b32, 0), c(b32, 3), c(b32, 0)), (b32, b32, b32, b32, a32, b32))
    assert out[3, 4, 3] == pytest.approx(exp, rel=2e-6)
    # spike = x-1 of (3,3,4): c4 sees three differences
    exp = mean(b32, (c(b32, 0), c(b32, 0), c(b32, 0), c(b32, 3)), (b32, b32, b32, b32, b32, a32))
    assert out[3, 3, 4] == pytest.approx(exp, rel=2e-6)
    # spike = x+1 of (3,3,2): one difference in c1, and a forward face value
    exp = mean(b32, (c(b32, 1), c(b32, 0), c(b32, 0), c(b32, 0)), (b32, b32, a32, b32, b32, b32))
    assert out[3, 3, 2] == pytest.approx(exp, rel=2e-6)


def test_cs3d_fp32_close_to_fp64():
    rng = np.random.default_rng(3)
    u = rng.random((12, 13, 14)).astype(np.float32)
    a = oracle.run(CS3D, [u.copy()], steps=4)[0]
    b = oracle.run(CS3D, [u.copy()], steps=4, dtype=np.float64)[0]
    assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


def test_cs3d_requires_a_wave_tile(otlib):
    ok = otlib.make_params(otlib.CS3D, n=64, steps=4, T=2, x=60, y=8, z=8, e=1, geometry="tile")
    d = otlib.derive(ok)
    assert d.threads_per_block == 256 and d.lds_bytes == 0
    bad = otlib.make_params(otlib.CS3D, n=64, steps=4, T=2, x=16, y=8, z=8, e=1, geometry="tile")
    with pytest.raises(otlib.OtError) as e:
        otlib.derive(bad)
    assert e.value.status == otlib.OT_EUNSUPPORTED and "wave-pipelined" in str(e.value)
