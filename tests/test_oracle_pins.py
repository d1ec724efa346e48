"""Pins of the oracle functions that restate generated kernel text rather than
a reference CPU loop: gradient-2d (whose CPU loop is also restated), gradient-3d,
TV update, Rician denoise, the 7-point Poisson relaxation with a right-hand side
and FDTD-2D. Runs on the CPU, no GPU needed.

Each stencil has one check function built from what the mathematics fixes,
without restating the oracle's own expression:
  * the tap set: perturbing one input value changes, after one step, exactly
    the output points whose stencil reads it (a wrong, shifted or swapped
    offset moves a point in or out of that set);
  * closed forms around a single spike on a constant field, where every
    coefficient reduces to u / sqrt(eps + k d^2) with k counted from which
    differences see the spike (pins eps, the number of terms, which
    coefficient takes which differences and which value each weight
    multiplies);
  * fixed points (discrete Poisson solutions in every axis), quadratic fields
    whose edge weights differ on the two sides of a point (pins which value is
    paired with which weight), and the Rician correction factor against the
    Bessel ratio I1(x)/I0(x) it approximates.
The second half plants a plausible mistake in a copy of the oracle's source
(a swapped tap, a wrong pairing, a dropped term, a wrong constant), builds it
with the same compiler flags and checks that the pins of that stencil fail.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.special import i0e, i1e

import oracle

J3D, P3D7, FDTD2D, G2D, G3D, RICIAN2D, TV2D = 4, 7, 8, 9, 10, 11, 12
F32 = lambda v: float(np.float32(v))  # noqa: E731  the float literal widened, as the FP64 oracle does


def _run64(L, st, rw, ro=(), coef=(0, 0, 0, 0), variant=0, steps=1):
    return oracle.run(st, [np.array(a, np.float64) for a in rw], [np.array(a, np.float64) for a in ro],
                      steps=steps, coef=coef, variant=variant, dtype=np.float64, library=L)


def _influence(L, st, rw, ro, src, p, coef=(0, 0, 0, 0), variant=0):
    """Per rw output field, the set of points whose value after one step moves
    when input `src` (("rw", k) or ("ro", k)) is perturbed at point p."""
    base = _run64(L, st, rw, ro, coef, variant)
    rw2 = [np.array(a, np.float64) for a in rw]
    ro2 = [np.array(a, np.float64) for a in ro]
    (rw2 if src[0] == "rw" else ro2)[src[1]][p] += 0.125
    pert = _run64(L, st, rw2, ro2, coef, variant)
    return [{tuple(int(i) for i in idx) for idx in np.argwhere(a != b)} for a, b in zip(base, pert)]


def _readers(p, taps):
    """Output points whose stencil reads input point p, given the stencil's
    taps as offsets (dz, dy, dx) of the values it reads."""
    return {(p[0] - t[0], p[1] - t[1], p[2] - t[2]) for t in taps}


STAR2 = [(0, 0, 0), (0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0)]
STAR3 = STAR2 + [(-1, 0, 0), (1, 0, 0)]
FACES3 = STAR3[1:]
# cdsc-tv-update-2d.cpp:266-285: tc, tr, ml, mc, mr, bl, bc
TV_TAPS = [(0, -1, 0), (0, -1, 1), (0, 0, -1), (0, 0, 0), (0, 0, 1), (0, 1, -1), (0, 1, 0)]
DIAMOND2 = [(0, dy, dx) for dy in range(-2, 3) for dx in range(-2, 3) if abs(dy) + abs(dx) <= 2]


def _close(got, expect, rel=1e-13):
    assert abs(got - expect) <= rel * abs(expect), (got, expect)


# --------------------------------------------------------------- gradient 2D/3D

def check_gradient(L, st):
    """gradient-2d.cpp:691-699, gradient-3d.cpp:283-307: c + 1/sqrt(1e-4 + sum (c-n)^2)."""
    dims = 2 if st == G2D else 3
    rng = np.random.default_rng(21)
    shape = (1, 9, 10) if dims == 2 else (8, 9, 10)
    u = 0.5 + rng.random(shape)
    p = (0, 4, 5) if dims == 2 else (4, 4, 5)
    assert _influence(L, st, [u], [], ("rw", 0), p)[0] == _readers(p, STAR2 if dims == 2 else STAR3)
    # a spike a on a constant b: the centre sees 2*dims differences, each face
    # neighbour one, every other point none
    a, b = 0.875, 0.25
    shape = (1, 9, 9) if dims == 2 else (9, 9, 9)
    s = (0, 4, 4) if dims == 2 else (4, 4, 4)
    u = np.full(shape, b)
    u[s] = a
    out = _run64(L, st, [u])[0]
    eps, d = F32(1e-4), (a - b) ** 2
    _close(out[s], a + 1 / math.sqrt(eps + 2 * dims * d))
    for t in (STAR2 if dims == 2 else STAR3)[1:]:
        _close(out[s[0] + t[0], s[1] + t[1], s[2] + t[2]], b + 1 / math.sqrt(eps + d))
    far = (0, 2, 2) if dims == 2 else (2, 2, 2)
    _close(out[far], b + 1 / math.sqrt(eps))


# ----------------------------------------------------------------------- TV

def check_tv(L):
    """cdsc-tv-update-2d.cpp:290-309 (c3 duplicates c1)."""
    rng = np.random.default_rng(22)
    shape = (1, 9, 10)
    u, V, F = 0.5 + rng.random(shape), 0.5 + rng.random(shape), rng.random(shape)
    p = (0, 4, 5)
    coef = (0.125, 0, 0, 0)
    assert _influence(L, TV2D, [u], [V, F], ("rw", 0), p, coef)[0] == _readers(p, TV_TAPS)
    assert _influence(L, TV2D, [u], [V, F], ("ro", 0), p, coef)[0] == {p}
    assert _influence(L, TV2D, [u], [V, F], ("ro", 1), p, coef)[0] == {p}
    # spike closed forms: every coefficient is u0 / sqrt(EPI1 + k d) / V, with k
    # the number of its two differences that involve the spike
    a, b, v0, f0, alpha = 0.875, 0.25, 0.75, 0.375, 0.125
    u = np.full((1, 11, 11), b)
    u[0, 5, 5] = a
    out = _run64(L, TV2D, [u], [np.full(u.shape, v0), np.full(u.shape, f0)], coef)[0]
    e1, d = F32(1e-8), (a - b) ** 2

    def w(u0, k):
        return u0 / math.sqrt(e1 + k * d) / v0

    def upd(c1, c2, c3, c4, bc=b, tc=b, mr=b, ml=b):
        return (alpha * f0 + c1 * bc + c2 * tc + c3 * mr + c4 * ml) / (alpha + c1 + c2 + c3 + c4)

    wa, wb0, wb1, wb2 = w(a, 1), w(b, 0), w(b, 1), w(b, 2)
    cases = {  # (dy, dx) of the point from the spike -> expected value
        (0, 0): upd(w(a, 2), wa, w(a, 2), wa),           # spike is mc: c1 = c3 see two differences
        (-1, 0): upd(wb1, wb0, wb1, wb0, bc=a),          # spike is bc (c1, c3 and the c1 weight)
        (0, -1): upd(wb1, wb0, wb1, wb0, mr=a),          # spike is mr (c1, c3 and the c3 weight)
        (1, 0): upd(wb0, wb2, wb0, wb0, tc=a),           # spike is tc (c2 twice, the c2 weight)
        (1, -1): upd(wb0, wb1, wb0, wb0),                # spike is tr (c2 once)
        (0, 1): upd(wb0, wb0, wb0, wb2, ml=a),           # spike is ml (c4 twice, the c4 weight)
        (-1, 1): upd(wb0, wb0, wb0, wb1),                # spike is bl (c4 once)
        (3, 3): upd(wb0, wb0, wb0, wb0),
    }
    for (dy, dx), expect in cases.items():
        _close(out[0, 5 + dy, 5 + dx], expect, rel=1e-12)


# ------------------------------------------------------------------- Rician

def _bessel_ratio(x):
    return i1e(x) / i0e(x)


def check_rician(L):
    """rician-denoise-2d.cpp:153-160 (constants), :271-281 (g, r), :292-327 (u)."""
    rng = np.random.default_rng(23)
    shape = (1, 11, 12)
    u = 0.5 + rng.random(shape)
    f = 0.5 + rng.random(shape)
    p = (0, 5, 6)
    assert _influence(L, RICIAN2D, [u], [f], ("rw", 0), p)[0] == _readers(p, DIAMOND2)
    assert _influence(L, RICIAN2D, [u], [f], ("ro", 0), p)[0] == {p}
    dt = 5.0
    gamma = F32(np.float32(0.065) / np.float32(0.0025))
    # with f = 0 the correction term vanishes and u' is the weighted mean
    # (u + dt sum u_n g_n) / (1 + dt (sum g_n + gamma)), each neighbour's value
    # weighted by its own g = 1/sqrt(EPSILON + sum of its squared differences).
    # On u = c0 + s q^2 along one axis q, g depends on q only and differs on the
    # two sides of every point, so a value paired with the other side's g shows.
    c0, s = 0.5, 0.01
    ny, nx = 12, 13
    eps = F32(1e-20)
    zero = np.zeros((1, ny, nx))

    def g_of(q):
        return 1 / math.sqrt(eps + s * s * ((2 * q + 1) ** 2 + (2 * q - 1) ** 2))

    def val(q):
        return c0 + s * q * q

    for axis in ("y", "x"):
        qq = np.arange(ny if axis == "y" else nx, dtype=np.float64)
        field = (val(qq)[None, :, None] * np.ones((1, 1, nx))) if axis == "y" else \
            (val(qq)[None, None, :] * np.ones((1, ny, 1)))
        out = _run64(L, RICIAN2D, [field], [zero])[0]
        for q in range(3, (ny if axis == "y" else nx) - 3):
            num = val(q) + dt * (val(q + 1) * g_of(q + 1) + val(q - 1) * g_of(q - 1) + 2 * val(q) * g_of(q))
            den = 1 + dt * (g_of(q + 1) + g_of(q - 1) + 2 * g_of(q) + gamma)
            got = out[0, q, 6] if axis == "y" else out[0, 6, q]
            _close(got, num / den, rel=1e-12)
    # the correction r(x), x = u f / SIGMA2, approximates I1(x)/I0(x) (the
    # Rician likelihood's Bessel ratio) to 7.2e-4 with the reference's
    # coefficients; a 1% change of any of them moves it by more than 2e-3. On a
    # checkerboard of 1 and 3 every g two points from the ring is 1/sqrt(16), so
    # r can be solved for from the update; f sweeps x over [0.01, 1000].
    ny, nx = 24, 24
    yy, xx = np.mgrid[0:ny, 0:nx]
    board = np.where((yy + xx) % 2 == 0, 1.0, 3.0)[None]
    xs = np.logspace(-2, 3, ny * nx).reshape(1, ny, nx)
    sigma2 = F32(0.0025)
    fld = xs * sigma2 / board
    out = _run64(L, RICIAN2D, [board], [fld])[0]
    g0 = 0.25
    den = 1 + dt * (4 * g0 + gamma)
    inner = (slice(None), slice(2, ny - 2), slice(2, nx - 2))
    other = 4.0 - board
    r = (out * den - board - dt * 4 * other * g0) / (dt * gamma * fld)
    err = np.abs(r[inner] - _bessel_ratio(xs[inner]))
    assert err.max() <= 1.0e-3, err.max()
    # the constants are the reference's float literals: SIGMA2 = 0.0025f (not
    # 0.05f * 0.05f) and GAMMA = LAMBDA / SIGMA2 evaluated in float; a point of
    # the checkerboard by hand in float32, in the kernel text's order
    f32 = np.float32
    one, three, dt32 = f32(1), f32(3), f32(5)
    fv = f32(0.0625)
    out32 = oracle.run(RICIAN2D, [board.astype(np.float32)], [np.full(board.shape, fv, np.float32)],
                       steps=1, library=L)[0]
    g = f32(0.25)
    rr = (one * fv) / f32(0.0025)
    rr = (rr * (f32(2.38944) + rr * (f32(0.950037) + rr))) / (f32(4.65314) + rr * (f32(2.57541) + rr * (
        f32(1.48937) + rr)))
    gm = f32(0.065) / f32(0.0025)
    num = one + dt32 * ((((three * g + three * g) + three * g) + three * g) + (gm * fv) * rr)
    hand = num / (one + dt32 * ((((g + g) + g) + g) + gm))
    assert out32[0, 6, 6] == hand, (out32[0, 6, 6], hand)


# ------------------------------------------------------------ Poisson 7-point

def check_p3d7(L):
    """u' = (sum of the six face neighbours - f) / 6."""
    rng = np.random.default_rng(24)
    shape = (8, 9, 10)
    u, f = rng.random(shape), rng.random(shape) - 0.5
    p = (4, 4, 5)
    assert _influence(L, P3D7, [u], [f], ("rw", 0), p)[0] == _readers(p, FACES3)
    assert _influence(L, P3D7, [u], [f], ("ro", 0), p)[0] == {p}
    # u = a x^2 + b y^2 + c z^2 has discrete Laplacian sum6 - 6u = 2(a + b + c):
    # with f equal to it u is a fixed point away from the ring, for each axis on
    # its own and for all three with different weights
    nz, ny, nx = 9, 10, 11
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for wa, wb, wc in ((1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 2, 3)):
        u = (wa * x * x + wb * y * y + wc * z * z).astype(np.float64)
        fv = np.full(u.shape, 2.0 * (wa + wb + wc))
        out = _run64(L, P3D7, [u], [fv])[0]
        core = (slice(1, -1),) * 3
        assert np.allclose(out[core], u[core], rtol=1e-14, atol=1e-12), (wa, wb, wc)


# --------------------------------------------------------------------- FDTD

def check_fdtd(L):
    """fdtd-2d.cpp:269-330: Ey' = Ey - 0.5 (Hz - Hz(y-1)), Ex' = Ex - 0.5 (Hz - Hz(x-1)),
    Hz' = Hz - 0.7 ((Ex'(x+1) - Ex') + Ey'(y+1) +/- Ey')."""
    rng = np.random.default_rng(25)
    shape = (1, 9, 10)
    fields = [rng.random(shape) - 0.5 for _ in range(3)]
    q = (0, 4, 5)
    X, Y = (0, 0, 1), (0, 1, 0)

    def sh(t, k=1):
        return (q[0] + k * t[0], q[1] + k * t[1], q[2] + k * t[2])

    for variant in (0, 1):
        ex, ey, hz = _influence(L, FDTD2D, fields, [], ("rw", 0), q, variant=variant)
        assert ex == {q} and ey == set() and hz == {q, sh(X, -1)}
        ex, ey, hz = _influence(L, FDTD2D, fields, [], ("rw", 1), q, variant=variant)
        assert ex == set() and ey == {q} and hz == {q, sh(Y, -1)}
        ex, ey, hz = _influence(L, FDTD2D, fields, [], ("rw", 2), q, variant=variant)
        assert ex == {q, sh(X)} and ey == {q, sh(Y)} and hz == {q, sh(X), sh(X, -1), sh(Y), sh(Y, -1)}
    # unit spikes in Ex and in Ey on zero fields: the Hz update sees them
    # through its x difference and its y term with the coefficient 0.7
    z = np.zeros((1, 9, 9))
    e = z.copy()
    e[0, 4, 4] = 1.0
    _, _, hz = _run64(L, FDTD2D, [e, z, z])
    assert hz[0, 4, 3] == -F32(0.7) and hz[0, 4, 4] == F32(0.7) and np.count_nonzero(hz) == 2
    for variant, sign in ((0, -1.0), (1, 1.0)):
        _, _, hz = _run64(L, FDTD2D, [z, e, z], variant=variant)
        assert hz[0, 3, 4] == -F32(0.7) and hz[0, 4, 4] == sign * F32(0.7) and np.count_nonzero(hz) == 2


CHECKS = {
    "gradient-2d": lambda L: check_gradient(L, G2D),
    "gradient-3d": lambda L: check_gradient(L, G3D),
    "tv-update-2d": check_tv,
    "rician-2d": check_rician,
    "poisson-3d-rhs": check_p3d7,
    "fdtd-2d": check_fdtd,
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_pins_hold_for_the_oracle(name):
    CHECKS[name](oracle.lib())


def test_rician_default_constants_equal_the_explicit_literals():
    rng = np.random.default_rng(26)
    u = rng.random((1, 14, 15)).astype(np.float32)
    f = rng.random((1, 14, 15)).astype(np.float32)
    a = oracle.run(RICIAN2D, [u.copy()], [f], steps=3)[0]
    b = oracle.run(RICIAN2D, [u.copy()], [f], steps=3, coef=[0.0025, 0.065, 5.0, 0])[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# -------------------------------------------------- planted mistakes must fail

ORACLE_DIR = os.path.dirname(oracle.LIB_PATH)
# (name, stencil check, text in ot_oracle_body.h, replacement)
MUTATIONS = [
    ("g2d-wrong-tap", "gradient-2d", "SQ(c - r) + SQ(c - t) + SQ(c - bt)", "SQ(c - r) + SQ(c - l) + SQ(c - bt)"),
    ("g2d-eps", "gradient-2d", "((REAL)0.0001f + SQ(c - l)", "((REAL)0.001f + SQ(c - l)"),
    ("g2d-dropped-term", "gradient-2d", "SQ(c - l) + SQ(c - r) + SQ(c - t)", "SQ(c - l) + SQ(c - t)"),
    ("g3d-wrong-z-tap", "gradient-3d", "SQ(c - a[i + sx]) + SQ(c - a[i - sxy])", "SQ(c - a[i + sx]) + SQ(c - a[i - sx - 1])"),
    ("g3d-dropped-term", "gradient-3d", "SQ(c - a[i + sx]) + SQ(c - a[i - sxy]) + SQ(c - a[i + sxy])",
     "SQ(c - a[i + sx]) + SQ(c - a[i - sxy])"),
    ("tv-tr-as-tl", "tv-update-2d", "SQ(mc - tc) + SQ(tr - tc)", "SQ(mc - tc) + SQ(a[i - sx - 1] - tc)"),
    ("tv-c2-weights-bc", "tv-update-2d", "c2 * tc", "c2 * bc"),
    ("tv-c3-not-c1", "tv-update-2d",
     "const REAL c3 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(bc - mc) + SQ(mr - mc)))",
     "const REAL c3 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(a[i + sx + 1] - mr) + SQ(mr - mc)))"),
    ("tv-c3-as-c2", "tv-update-2d",
     "const REAL c3 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(bc - mc) + SQ(mr - mc)))",
     "const REAL c3 = mc * ((REAL)1.0f / RS_ARG((REAL)1e-8f + SQ(mc - tc) + SQ(tr - tc)))"),
    ("tv-alpha-times-u", "tv-update-2d", "(alpha * F[i]", "(alpha * a[i]"),
    ("tv-eps", "tv-update-2d", "(REAL)1e-8f + SQ(bl - ml)", "(REAL)1e-6f + SQ(bl - ml)"),
    ("rician-bottom-with-top-g", "rician-2d", "a[i + sx] * gb", "a[i + sx] * gt"),
    ("rician-left-right-g-swapped", "rician-2d", "a[i + 1] * gr + a[i - 1] * gl", "a[i + 1] * gl + a[i - 1] * gr"),
    ("rician-poly-digit", "rician-2d", "(REAL)2.38944f", "(REAL)2.83944f"),
    ("rician-g-diagonal-tap", "rician-2d", "SQ(c - a[i + sx]) +\n", "SQ(c - a[i + sx + 1]) +\n"),
    ("rician-r-without-f", "rician-2d", "REAL r = c * f / sigma2;", "REAL r = c / sigma2;"),
    ("rician-sigma-squared-in-float", "rician-2d", "(REAL)coef[0] : (REAL)0.0025f", "(REAL)coef[0] : (REAL)(0.05f * 0.05f)"),
    ("rician-gamma-missing-from-denominator", "rician-2d", "gb + gt + gamma))", "gb + gt))"),
    ("p3d7-y-tap-as-z", "poisson-3d-rhs", "a[i - sx] + a[i + sx] + a[i - sxy] + a[i + sxy] - f[i]",
     "a[i - sx] + a[i + sxy] + a[i - sxy] + a[i + sxy] - f[i]"),
    ("p3d7-rhs-sign", "poisson-3d-rhs", "- f[i]) / (REAL)6.0f", "+ f[i]) / (REAL)6.0f"),
    ("fdtd-ey-from-hz-below", "fdtd-2d", "(h - hz[i - sx])", "(h - hz[i + sx])"),
    ("fdtd-hz-ey-above", "fdtd-2d", "ey[i + sx] + ey[i])", "ey[i - sx] + ey[i])"),
    ("fdtd-default-takes-the-curl-sign", "fdtd-2d", "ex[i + 1] - ex[i] + ey[i + sx] + ey[i])",
     "ex[i + 1] - ex[i] + ey[i + sx] - ey[i])"),
    ("fdtd-hz-coefficient", "fdtd-2d", "hz[i] - (REAL)0.7f * (ex[i + 1] - ex[i] + ey[i + sx] + ey[i])",
     "hz[i] - (REAL)0.75f * (ex[i + 1] - ex[i] + ey[i + sx] + ey[i])"),
]


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


def _build_mutant(dirpath, name, old, new):
    body = open(os.path.join(ORACLE_DIR, "ot_oracle_body.h")).read()
    assert body.count(old) == 1, f"{name}: the planted text must occur exactly once"
    d = os.path.join(str(dirpath), name)
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ORACLE_DIR, "ot_oracle.c"), d)
    with open(os.path.join(d, "ot_oracle_body.h"), "w") as fh:
        fh.write(body.replace(old, new))
    so = os.path.join(d, "libmutant.so")
    # the flags of oracle/Makefile
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-ffp-contract=off", "-shared", "-o", so,
                    os.path.join(d, "ot_oracle.c"), "-lm"], check=True, capture_output=True)
    return oracle.load(so)


@pytest.mark.parametrize("name,check,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_fails_a_pin(mutant_dir, name, check, old, new):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    L = _build_mutant(mutant_dir, name, old, new)
    with pytest.raises(AssertionError):
        CHECKS[check](L)
