"""The kernels' rsqrt (hardware estimates + one-ulp square-root correction +
Newton/Markstein reciprocal) equals the IEEE expression 1.0f / sqrtf(x) bit for
bit on every float the stencils can pass it: all of [1e-20, +inf] (the
smallest argument is Rician's 1e-20 + sum of squares), plus NaN inputs."""
import ctypes as C
import struct

import pytest

import otlib

pytestmark = pytest.mark.gpu


def _bits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


@pytest.mark.parametrize("lo,hi", [(_bits(1e-20), 0x7F800000), (0x7F800001, 0x7FFFFFFF)],
                         ids=["1e-20..inf", "nan"])
def test_rsq_matches_ieee_exhaustively(otlib, lo, hi):
    bad = C.c_uint64(0)
    first = C.c_uint32(0)
    otlib.check(otlib.lib().ot_selftest_rsq(lo, hi, C.byref(bad), C.byref(first)))
    print(f"rsq check {lo:#010x}..{hi:#010x}: {hi - lo + 1} inputs, {bad.value} differ, first {first.value:#010x}",
          flush=True)
    assert bad.value == 0, hex(first.value)
