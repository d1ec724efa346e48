"""Pins of the CPU oracle (runs on the CPU, no GPU needed).

The reference ships no golden vectors for the numerics of this path (its only
check is the drivers' --verify against inline CPU loops), so the oracle is
pinned by what the mathematics fixes, chosen so that a dropped term, a wrong
sign, a shifted index or a wrong coefficient fails at least one check:
  * single-delta responses of the linear stencils equal c^t times integer
    path counts obtained by exact integer convolution;
  * integer-valued Poisson-2D/3D iterates are exact in FP32;
  * constant fields, fixed points and hand-derived one-step FDTD values;
  * linearity of the FP64 oracle for the linear stencils;
  * the seeded glibc init, restated independently by the oracle and the engine.
"""
import math

import numpy as np
import pytest

import oracle

J1D3, J1D5, J1D7, J2D, J3D, P2D, P3D27, P3D7, FDTD2D, G2D, G3D, RICIAN2D, TV2D = range(13)


def _conv_power(kernel: np.ndarray, t: int, shape) -> np.ndarray:
    """Exact integer response of a delta after t applications of `kernel`
    (object arrays of Python ints; kernel axes ordered (z, y, x))."""
    out = np.zeros(shape, dtype=object)
    c = tuple(s // 2 for s in shape)
    out[c] = 1
    r = [k // 2 for k in kernel.shape]
    for _ in range(t):
        new = np.zeros(shape, dtype=object)
        for idx in np.ndindex(kernel.shape):
            w = int(kernel[idx])
            if w == 0:
                continue
            shift = [idx[a] - r[a] for a in range(3)]
            new += w * np.roll(out, shift=shift, axis=(0, 1, 2))
        out = new
    return out


def _delta(shape):
    u = np.zeros(shape, dtype=np.float32)
    u[tuple(s // 2 for s in shape)] = 1.0
    return u


@pytest.mark.parametrize("st,kernel,coef,t", [
    (J1D3, np.ones((1, 1, 3), int), 0.333, 6),
    (J1D5, np.ones((1, 1, 5), int), 0.2, 5),
    (J1D7, np.ones((1, 1, 7), int), 0.2, 4),
    (J2D, np.array([[[0, 1, 0], [1, 1, 1], [0, 1, 0]]]), 0.2, 6),
    (J3D, None, 0.143, 4),
])
def test_jacobi_delta_closed_form(st, kernel, coef, t):
    if st == J3D:
        kernel = np.zeros((3, 3, 3), int)
        kernel[1, 1, :] = 1
        kernel[1, :, 1] = 1
        kernel[:, 1, 1] = 1
    dims = oracle.DIMS[st]
    shape = {1: (1, 1, 61), 2: (1, 31, 33), 3: (21, 19, 23)}[dims]
    counts = _conv_power(kernel, t, shape).astype(np.float64)
    cf = float(np.float32(coef))
    expected = counts * cf ** t
    got64 = oracle.run(st, [_delta(shape)], steps=t, dtype=np.float64)[0]
    assert np.allclose(got64, expected, rtol=1e-12, atol=1e-300)
    got32 = oracle.run(st, [_delta(shape)], steps=t)[0]
    assert np.allclose(got32, expected, rtol=4e-6, atol=1e-30)
    # the response is confined to radius*t (no spurious taps)
    assert np.count_nonzero(got64) == np.count_nonzero(counts)


def test_poisson2d_integer_iterates_exact():
    k = -np.ones((1, 3, 3), int)
    k[0, 1, 1] = 8
    shape = (1, 25, 27)
    counts = _conv_power(k, 4, shape)
    got = oracle.run(P2D, [_delta(shape)], steps=4)[0]
    assert np.array_equal(got, counts.astype(np.float32))


def test_poisson3d27_integer_iterates_exact():
    k = -np.ones((3, 3, 3), int)
    k[1, 1, 1] = 26
    shape = (15, 15, 15)
    counts = _conv_power(k, 2, shape)
    got = oracle.run(P3D27, [_delta(shape)], steps=2)[0]
    assert np.array_equal(got, counts.astype(np.float32))


def test_poisson_constant_field_interior_zero():
    u = np.full((1, 20, 22), 3.0, np.float32)
    out = oracle.run(P2D, [u], steps=1)[0]
    assert np.all(out[0, 1:-1, 1:-1] == 0)
    assert np.all(out[0, 0, 1:-1] == 9.0)  # 8*3 - 5*3: three neighbours lie in the zero ring
    u3 = np.full((12, 12, 12), 2.0, np.float32)
    out3 = oracle.run(P3D27, [u3], steps=1)[0]
    assert np.all(out3[1:-1, 1:-1, 1:-1] == 0)


def test_poisson3d7_discrete_solution_is_fixed_point():
    # u = x^2 has discrete Laplacian 2, so with f = 2 it is a fixed point of
    # u' = (sum6 - f)/6 wherever the stencil does not touch the boundary.
    nz, ny, nx = 9, 9, 11
    x = np.arange(nx, dtype=np.float64)
    u = np.broadcast_to(x * x, (nz, ny, nx)).astype(np.float64).copy()
    f = np.full_like(u, 2.0)
    out = oracle.run(P3D7, [u], [f], steps=1, dtype=np.float64)[0]
    assert np.allclose(out[1:-1, 1:-1, 1:-1], u[1:-1, 1:-1, 1:-1], rtol=1e-14, atol=1e-12)
    # and zero u with constant f gives -f/6 in the interior
    out0 = oracle.run(P3D7, [np.zeros((5, 5, 5), np.float32)], [np.full((5, 5, 5), 3.0, np.float32)],
                      steps=1)[0]
    assert np.all(out0 == np.float32(-3.0) / np.float32(6.0))


@pytest.mark.parametrize("st", [G2D, G3D])
def test_gradient_constant_field_adds_inverse_sqrt_eps(st):
    dims = oracle.DIMS[st]
    shape = (1, 24, 26) if dims == 2 else (14, 15, 16)
    t = 3
    c = np.float32(0.25)
    out = oracle.run(st, [np.full(shape, c, np.float32)], steps=t)[0]
    step = np.float32(1.0) / np.sqrt(np.float32(0.0001))
    expect = c
    for _ in range(t):
        expect = np.float32(expect + step)
    core = tuple(slice(t, s - t) if s > 1 else slice(None) for s in shape)
    assert np.all(out[core] == expect)
    # and the double-sqrt overload variant of gradient-2d agrees to FP32 rounding
    if st == G2D:
        rng = np.random.default_rng(1)
        u = rng.random(shape, dtype=np.float32)
        a = oracle.run(G2D, [u.copy()], steps=4)[0]
        b = oracle.g2d_double_sqrt(u.copy(), 4)
        assert np.allclose(a, b, rtol=1e-6)


def test_tv_fixed_point():
    shape = (1, 20, 21)
    u = np.full(shape, 0.7, np.float32)
    out = oracle.run(TV2D, [u.copy()], [np.ones(shape, np.float32), u.copy()], steps=4,
                     coef=[0.1, 0, 0, 0])[0]
    assert np.allclose(out[0, 4:-4, 4:-4], 0.7, rtol=1e-6)


def test_rician_near_fixed_point():
    shape = (1, 20, 21)
    u = np.full(shape, 0.6, np.float32)
    out = oracle.run(RICIAN2D, [u.copy()], [u.copy()], steps=3)[0]
    assert np.allclose(out[0, 3:-3, 3:-3], 0.6, rtol=1e-6)


def test_rician_single_step_hand_value():
    # one interior point with neighbours all equal: g = 1/sqrt(1e-20 + 4 d^2)
    shape = (1, 5, 5)
    u = np.zeros(shape, np.float64)
    u[0, 1:4, 1:4] = 1.0
    u[0, 2, 2] = 2.0
    f = np.full(shape, 1.0, np.float64)
    out = oracle.run(RICIAN2D, [u.copy()], [f], steps=1, dtype=np.float64)[0]
    # centre: its four neighbours have value 1, each neighbour's g uses its own
    # 4 neighbours: (1-2)^2 + 2*(1-1)^2 + (1-0)^2 = 2 -> g_n = 1/sqrt(2)
    sigma2 = float(np.float32(0.0025))
    gamma = float(np.float32(np.float32(0.065) / np.float32(0.0025)))
    dt = 5.0
    r = 2.0 * 1.0 / sigma2
    r = (r * (float(np.float32(2.38944)) + r * (float(np.float32(0.950037)) + r))) / (
        float(np.float32(4.65314)) + r * (float(np.float32(2.57541)) + r * (float(np.float32(1.48937)) + r)))
    gn = 1.0 / math.sqrt(float(np.float32(1e-20)) + 2.0)
    expect = (2.0 + dt * (4 * 1.0 * gn + gamma * 1.0 * r)) / (1.0 + dt * (4 * gn + gamma))
    assert abs(out[0, 2, 2] - expect) < 1e-12 * abs(expect)


def test_fdtd_one_step_hand_values():
    shape = (1, 9, 9)
    hz = np.zeros(shape, np.float32)
    hz[0, 4, 4] = 1.0
    z = np.zeros(shape, np.float32)
    ex, ey, hz1 = oracle.run(FDTD2D, [z.copy(), z.copy(), hz.copy()], steps=1)
    assert ey[0, 4, 4] == -0.5 and ey[0, 5, 4] == 0.5
    assert ex[0, 4, 4] == -0.5 and ex[0, 4, 5] == 0.5
    assert hz1[0, 4, 4] == np.float32(1.0) - np.float32(0.7) * np.float32(1.0)  # reference "+ey" sign
    ex, ey, hzp = oracle.run(FDTD2D, [z.copy(), z.copy(), hz.copy()], steps=1, variant=1)
    assert hzp[0, 4, 4] == np.float32(1.0) - np.float32(0.7) * np.float32(2.0)  # physical curl
    # zero fields stay zero
    out = oracle.run(FDTD2D, [z.copy(), z.copy(), z.copy()], steps=5)
    assert all(np.all(a == 0) for a in out)


@pytest.mark.parametrize("st", [J2D, J3D, P2D, P3D27, FDTD2D, J1D5])
def test_linearity_fp64(st):
    dims = oracle.DIMS[st]
    shape = {1: (1, 1, 40), 2: (1, 17, 19), 3: (9, 10, 11)}[dims]
    rng = np.random.default_rng(st)
    nf = oracle.NUM_RW.get(st, 1)
    x = [rng.random(shape) for _ in range(nf)]
    y = [rng.random(shape) for _ in range(nf)]
    a, b = 1.5, -0.75
    rx = oracle.run(st, [v.copy() for v in x], steps=3, dtype=np.float64)
    ry = oracle.run(st, [v.copy() for v in y], steps=3, dtype=np.float64)
    rz = oracle.run(st, [a * u + b * v for u, v in zip(x, y)], steps=3, dtype=np.float64)
    for p, q, r in zip(rx, ry, rz):
        assert np.allclose(r, a * p + b * q, rtol=1e-12, atol=1e-12)


def test_fp32_vs_fp64_small_error():
    rng = np.random.default_rng(3)
    u = rng.random((12, 13, 14), dtype=np.float32)
    a = oracle.run(J3D, [u.copy()], steps=8)[0]
    b = oracle.run(J3D, [u.astype(np.float64)], steps=8, dtype=np.float64)[0]
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-5


def test_compare_metrics():
    h = np.array([[[3.0, 4.0]]], np.float32)
    d = np.array([[[3.0, 4.5]]], np.float32)
    m = oracle.compare(h, d)
    assert m["error_norm"] == pytest.approx(0.5)
    assert m["ref_norm"] == pytest.approx(5.0)
    assert m["max_rel"] == pytest.approx(0.125)
    assert m["correctness"] == "FAILED"
    assert oracle.compare(h, h)["bitwise_equal"]


def test_reference_init_matches_engine_generator(otlib):
    """Two independent restatements of the drivers' srand(123456)/rand() init
    (poisson-2d.cpp:611-621) produce the same padded arrays."""
    for st, n, T in [(otlib.J2D, 70, 3), (otlib.J3D, 50, 2), (otlib.J1D5, 300, 2)]:
        p = otlib.make_params(st, n=n, T=T, x=16, y=8, z=8, geometry="reference")
        d = otlib.derive(p)
        eng = otlib.fill_reference_uniform(d, 123456)
        sz, sy, sx = otlib.layout_shape(d)
        orc = oracle.reference_init(d.dims, (sz, sy, sx), d.padding, 0, 123456)
        assert np.array_equal(eng, orc)
    # first values of glibc's sequence for srand(123456), scaled as the drivers do
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(123456)
    first = np.float32(libc.rand()) / (np.float32(2147483647) + np.float32(1.0))
    p = otlib.make_params(otlib.J2D, n=70, T=3, x=16, y=16, geometry="reference")
    d = otlib.derive(p)
    eng = otlib.fill_reference_uniform(d, 123456)
    assert eng[otlib.interior_slices(d)][0, 0, 0] == first
