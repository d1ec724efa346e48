"""ot_derive against the reference's own recorded geometry (CPU only).

tests/golden/geometry_c2050.json holds rows of the reference's committed
Tesla C2050 sweeps (scripts/c2050/*.dat, written by scripts/utils.py from the
drivers' stdout): driver inputs and the derived geometry/per-block keys each
driver printed. They are known-answer vectors for GeneratorParams::computeDerived
and ProgramGenerator::printProgramParameters, which ot_derive restates in
reference geometry.
"""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
VECTORS = json.load(open(os.path.join(HERE, "golden", "geometry_c2050.json")))

def test_golden_vectors_present():
    stencils = {r["stencil"] for r in VECTORS}
    assert len(VECTORS) > 500
    assert stencils == {"jacobi-2d", "jacobi-3d", "poisson-2d", "poisson-3d", "gradient-2d",
                        "gradient-3d", "fdtd-2d", "rician-denoise-2d", "cdsc-tv-update-2d"}


@pytest.mark.parametrize("stencil", sorted({r["stencil"] for r in VECTORS}))
def test_reference_geometry_matches_c2050_sweeps(otlib, stencil):
    import ctypes as C
    checked = 0
    for r in (v for v in VECTORS if v["stencil"] == stencil):
        p = otlib.make_params(r["stencil"], n=r["problem_size"], x=r["Block Size X"], y=r["Block Size Y"],
                              z=r["Block Size Z"], e=r["Elements/Thread"], T=r["Time Tile Size"],
                              steps=r["Time Steps"], geometry="reference")
        d = otlib.Derived()
        st = otlib.lib().ot_derive(C.byref(p), C.byref(d))
        # the reference ran these configurations, so they are valid geometry;
        # this engine may only refuse them for its own resource limits
        assert st in (otlib.OT_OK, otlib.OT_ERESOURCE), (r, otlib.lib().ot_last_error())
        steps_scaled = r["stencil"] in ("poisson-2d", "cdsc-tv-update-2d")
        got = {
            "Num Blocks X": d.num_blocks[0], "Num Blocks Y": d.num_blocks[1], "Num Blocks Z": d.num_blocks[2],
            "Padded Size": d.padded_size[0], "Real Size": d.real_size[0], "Shared Size": d.shared_size,
            "Real Per Block X": d.real_per_block[0], "Real Per Block Y": d.real_per_block[1],
            "Total FP/Block": d.total_fp_per_block, "Useful FP/Block": d.useful_fp_per_block,
            "Total FP": d.total_fp_steps if steps_scaled else d.total_fp_per_block,
            "Useful FP": d.useful_fp_steps if steps_scaled else d.useful_fp_per_block,
            "Useful Ratio": d.useful_ratio, "Global Loads/Block": d.global_loads_per_block,
            "Global Stores/Block": d.global_stores_per_block, "Shared Loads/Block": d.shared_loads_per_block,
            "Shared Stores/Block": d.shared_stores_per_block,
        }
        for k, v in r.items():
            if k in got:
                assert got[k] == pytest.approx(v, rel=1e-4, abs=1e-6), (k, r)
                checked += 1
    assert checked > 100


def test_reference_geometry_examples(otlib):
    # jacobi-2d: x32 y4 e1 T1, n=6000 -> gcd(32, 4)=4, realSize 6000, 187 x 1500 blocks
    d = otlib.derive(otlib.make_params("jacobi-2d", n=6000, x=32, y=4, e=1, T=1, geometry="reference"))
    assert (d.real_size[0], d.num_blocks[0], d.num_blocks[1], d.padded_size[0]) == (6000, 187, 1500, 6002)
    assert d.grid_blocks[0] == 188  # this engine also computes the columns the reference leaves out
    # Jacobi-2D 1024^2, 64 steps, T=4, 16x16 in reference geometry: 1020^2, 16 launches
    d = otlib.derive(otlib.make_params("jacobi-2d", n=1024, x=16, y=16, e=1, T=4, t=64, geometry="reference"))
    assert (d.real_size[0], d.num_blocks[0], d.launches, d.padding) == (1020, 102, 16, 4)
    # jacobi-1d adds 109 to the padded size
    d = otlib.derive(otlib.make_params("jacobi-1d", n=1024, x=16, e=1, T=1, geometry="reference"))
    assert d.padded_size[0] == 1024 + 2 + 109
    # 3D lcm truncation: 8x8x4 at T=2 -> rpb 6,6,2 -> lcm 6 -> 252
    d = otlib.derive(otlib.make_params("jacobi-3d", n=256, x=8, y=8, z=4, e=1, T=2, geometry="reference"))
    assert d.real_size[0] == 252


def test_reference_geometry_invalid(otlib):
    with pytest.raises(otlib.OtError) as ex:
        otlib.derive(otlib.make_params("jacobi-3d", n=256, x=8, y=8, z=4, e=1, T=3, geometry="reference"))
    assert ex.value.status == otlib.OT_EINVAL and "Consistency" in str(ex.value)


@pytest.mark.parametrize("stencil", ["jacobi-1d", "jacobi-2d", "jacobi-3d", "rician-denoise-2d", "fdtd-2d"])
def test_tile_geometry_covers_domain(otlib, stencil):
    for n, T, bx, by, bz, e in [(100, 1, 16, 16, 8, 1), (257, 4, 32, 8, 16, 2), (64, 8, 8, 4, 4, 1)]:
        p = otlib.make_params(stencil, n=n, T=T, x=bx, y=by, z=bz, e=e, t=10)
        try:
            d = otlib.derive(p)
        except otlib.OtError as ex:
            assert ex.status == otlib.OT_ERESOURCE
            continue
        for a in range(d.dims):
            assert d.grid_blocks[a] * d.tile[a] >= n
            assert (d.grid_blocks[a] - 1) * d.tile[a] < n
        assert d.halo == d.radius * T
        assert d.launches == -(-10 // T)
        assert d.padding == 0 and d.origin == 0
        if d.dims > 1:
            assert d.stride_y % 16 == 0


def test_resource_limits(otlib):
    with pytest.raises(otlib.OtError) as ex:
        otlib.derive(otlib.make_params("jacobi-2d", n=512, x=64, y=64, e=1, T=4))
    assert ex.value.status == otlib.OT_ERESOURCE
    with pytest.raises(otlib.OtError) as ex:
        otlib.derive(otlib.make_params("jacobi-2d", n=512, x=16, y=16, e=1, T=9))
    assert ex.value.status == otlib.OT_EUNSUPPORTED
    with pytest.raises(otlib.OtError) as ex:
        otlib.derive(otlib.make_params("jacobi-2d", n=512, x=0, y=16, e=1, T=1))
    assert ex.value.status == otlib.OT_EINVAL
