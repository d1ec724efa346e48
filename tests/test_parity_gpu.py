"""GPU parity of every stencil against the CPU oracle (run on the MI355X box).

Each case runs the engine's overlapped time-tiled kernels through the C ABI and
the FP32 oracle on the same seeded inputs, and requires bit-identical results
(the kernels evaluate each formula in the reference's association order with FP
contraction off) — which implies the 1e-5 max-relative bar. Cases cover every
stencil, T from 1 to 8, several tile shapes and elements per thread, domains that
are not multiples of the tile (partial edge tiles), a time-step count that is not
a multiple of T (shortened last launch), both geometries, and graph replay.
"""
import numpy as np
import pytest

import oracle
import otlib
from helpers import check_against_oracle, domain_shape, random_inputs, run_gpu, MAX_REL

pytestmark = pytest.mark.gpu


def _params(st, n3, steps, T, block, ept, geometry="tile", **kw):
    nx, ny, nz = n3
    return otlib.make_params(st, nx=nx, ny=ny, nz=nz, problem_size=nx, steps=steps, T=T,
                             x=block[0], y=block[1], z=block[2], e=ept, geometry=geometry, **kw)


CASES_1D = [
    (otlib.J1D3, (1000, 1, 1), 13, 1, (64, 1, 1), 1),
    (otlib.J1D3, (1000, 1, 1), 13, 4, (64, 1, 1), 3),
    (otlib.J1D3, (4097, 1, 1), 17, 8, (128, 1, 1), 2),
    (otlib.J1D5, (777, 1, 1), 11, 3, (64, 1, 1), 2),
    (otlib.J1D7, (777, 1, 1), 11, 5, (32, 1, 1), 4),
]
CASES_2D = [
    (otlib.J2D, (100, 77, 1), 9, 1, (16, 16, 1), 1),
    (otlib.J2D, (100, 77, 1), 9, 4, (16, 16, 1), 1),
    (otlib.J2D, (130, 70, 1), 11, 3, (32, 4, 1), 3),
    (otlib.J2D, (64, 64, 1), 16, 8, (8, 8, 1), 2),
    (otlib.P2D, (90, 61, 1), 7, 3, (16, 8, 1), 2),
    (otlib.G2D, (90, 61, 1), 7, 2, (16, 16, 1), 1),
    (otlib.G2D, (90, 61, 1), 7, 4, (32, 8, 1), 4),
    (otlib.TV2D, (90, 61, 1), 9, 4, (16, 16, 1), 2),
    (otlib.RICIAN2D, (90, 61, 1), 9, 1, (16, 16, 1), 1),
    (otlib.RICIAN2D, (90, 61, 1), 9, 4, (16, 8, 1), 2),
    (otlib.FDTD2D, (90, 61, 1), 9, 1, (16, 16, 1), 1),
    (otlib.FDTD2D, (90, 61, 1), 9, 3, (16, 8, 1), 3),
    (otlib.FDTD2D, (64, 64, 1), 16, 8, (16, 16, 1), 1),
    # wave-per-tile kernels: x tile = 64 - 2*radius*T fills a wavefront
    (otlib.J2D, (130, 77, 1), 9, 2, (60, 12, 1), 1),
    (otlib.J2D, (100, 90, 1), 11, 4, (56, 24, 1), 1),
    (otlib.J2D, (150, 140, 1), 19, 8, (48, 48, 1), 1),
    (otlib.J2D, (70, 45, 1), 5, 1, (62, 14, 1), 1),
    (otlib.P2D, (90, 61, 1), 7, 3, (58, 26, 1), 1),
    (otlib.G2D, (90, 61, 1), 9, 4, (56, 8, 1), 1),
    (otlib.TV2D, (90, 61, 1), 7, 2, (60, 28, 1), 1),
    (otlib.TV2D, (100, 70, 1), 11, 5, (54, 6, 1), 1),
    (otlib.FDTD2D, (90, 61, 1), 8, 3, (58, 10, 1), 1),
    (otlib.FDTD2D, (130, 70, 1), 13, 6, (52, 20, 1), 1),
    (otlib.FDTD2D, (70, 40, 1), 5, 1, (62, 30, 1), 1),
    (otlib.RICIAN2D, (90, 61, 1), 7, 2, (56, 8, 1), 1),
    (otlib.RICIAN2D, (110, 70, 1), 9, 4, (48, 16, 1), 1),
    (otlib.RICIAN2D, (100, 50, 1), 7, 3, (52, 20, 1), 1),
    # 24 footprint rows
    (otlib.J2D, (100, 77, 1), 9, 4, (56, 16, 1), 1),
    (otlib.P2D, (90, 61, 1), 7, 2, (60, 20, 1), 1),
    (otlib.G2D, (90, 61, 1), 9, 3, (58, 18, 1), 1),
    (otlib.TV2D, (90, 61, 1), 7, 5, (54, 14, 1), 1),
    (otlib.FDTD2D, (130, 70, 1), 15, 7, (50, 10, 1), 1),
    (otlib.FDTD2D, (90, 61, 1), 8, 8, (48, 8, 1), 1),
    (otlib.RICIAN2D, (90, 61, 1), 7, 3, (52, 12, 1), 1),
    (otlib.RICIAN2D, (90, 61, 1), 9, 5, (44, 4, 1), 1),
    # row-streaming kernels: the same x tiles, y tile = streamed chunk of >= 64 rows
    (otlib.J2D, (200, 150, 1), 9, 2, (60, 64, 1), 1),
    (otlib.J2D, (130, 230, 1), 19, 8, (48, 100, 1), 1),
    (otlib.J2D, (70, 40, 1), 5, 1, (62, 64, 1), 1),
    (otlib.P2D, (130, 140, 1), 7, 3, (58, 70, 1), 1),
    (otlib.G2D, (90, 130, 1), 9, 4, (56, 64, 1), 1),
    (otlib.TV2D, (100, 150, 1), 7, 3, (58, 80, 1), 1),
    (otlib.TV2D, (120, 70, 1), 11, 5, (54, 64, 1), 1),
    (otlib.FDTD2D, (130, 170, 1), 8, 3, (58, 64, 1), 1),
    (otlib.FDTD2D, (110, 210, 1), 13, 6, (52, 100, 1), 1),
    (otlib.FDTD2D, (70, 90, 1), 5, 1, (62, 64, 1), 1),
    (otlib.RICIAN2D, (110, 150, 1), 7, 2, (56, 64, 1), 1),
    (otlib.RICIAN2D, (100, 200, 1), 9, 4, (48, 96, 1), 1),
    (otlib.RICIAN2D, (90, 64, 1), 7, 3, (52, 64, 1), 1),
]
CASES_3D = [
    (otlib.J3D, (40, 37, 35), 7, 1, (16, 8, 8), 1),
    (otlib.J3D, (40, 37, 35), 7, 2, (16, 8, 8), 2),
    (otlib.J3D, (40, 37, 35), 7, 4, (8, 8, 4), 4),
    (otlib.J3D, (33, 34, 35), 11, 8, (16, 16, 8), 1),
    (otlib.J3D, (32, 32, 32), 9, 3, (32, 16, 8), 2),
    (otlib.G3D, (36, 30, 29), 6, 3, (16, 8, 8), 2),
    (otlib.P3D7, (36, 30, 29), 7, 4, (16, 16, 8), 1),
    (otlib.P3D27, (36, 30, 29), 5, 2, (16, 8, 8), 1),
    (otlib.P3D27, (36, 30, 29), 5, 3, (8, 8, 4), 2),
    # wave-pipelined kernels: x tile = 64 - 2T fills a wavefront
    (otlib.J3D, (70, 37, 35), 7, 2, (60, 12, 8), 1),
    (otlib.J3D, (130, 40, 33), 9, 3, (58, 10, 16), 1),
    (otlib.J3D, (61, 29, 40), 11, 4, (56, 4, 16), 1),
    (otlib.J3D, (64, 30, 31), 7, 1, (62, 30, 8), 1),
    (otlib.G3D, (64, 30, 29), 6, 1, (62, 6, 8), 1),
    (otlib.G3D, (90, 33, 20), 8, 3, (58, 2, 8), 1),
    (otlib.P3D7, (70, 25, 29), 9, 4, (56, 4, 8), 1),
    (otlib.P3D27, (70, 25, 29), 5, 2, (60, 8, 8), 1),
    (otlib.P3D27, (66, 21, 17), 7, 3, (58, 6, 4), 1),
    (otlib.J3D, (70, 30, 30), 12, 5, (54, 2, 8), 1),
    # chunks away from both z faces: their middle iterations take the
    # unchecked steady-state blocks (T, T and a shortened last launch)
    (otlib.J3D, (70, 20, 60), 7, 3, (58, 6, 16), 1),
    (otlib.G3D, (64, 20, 50), 5, 2, (60, 8, 12), 1),
    (otlib.J3D, (62, 12, 50), 3, 1, (62, 6, 16), 1),
    # tiles whose lanes and rows all lie inside the domain in x and y (the tile
    # body without the per-point mask), in checked and steady-state iterations
    (otlib.J3D, (200, 40, 60), 7, 3, (58, 6, 16), 1),
    (otlib.J3D, (190, 44, 40), 5, 1, (62, 6, 8), 1),
    (otlib.G3D, (200, 40, 50), 5, 2, (60, 8, 12), 1),
    (otlib.P3D7, (200, 40, 60), 7, 2, (60, 4, 16), 1),
    # cs-3d (wave-pipelined kernel only): boundary tiles, a shortened last
    # launch, tiles inside the domain in x and y with steady-state blocks
    (otlib.CS3D, (70, 30, 29), 5, 2, (60, 8, 8), 1),
    (otlib.CS3D, (64, 30, 31), 7, 1, (62, 10, 8), 1),
    (otlib.CS3D, (130, 40, 33), 8, 3, (58, 10, 16), 1),
    (otlib.CS3D, (200, 40, 60), 7, 3, (58, 6, 16), 1),
    (otlib.CS3D, (61, 29, 40), 9, 4, (56, 4, 16), 1),
]


def _ids(cases):
    return [f"{otlib.STENCILS[c[0]]}-n{'x'.join(map(str, c[1]))}-t{c[2]}-T{c[3]}-b{'x'.join(map(str, c[4]))}-e{c[5]}"
            for c in cases]


@pytest.mark.parametrize("case", CASES_1D + CASES_2D + CASES_3D, ids=_ids(CASES_1D + CASES_2D + CASES_3D))
def test_tile_geometry_matches_oracle(otlib, case):
    st, n3, steps, T, block, ept = case
    p = _params(st, n3, steps, T, block, ept)
    d = otlib.derive(p)
    rw, ro = random_inputs(st, domain_shape(d), seed=st * 100 + T)
    print(f"case {otlib.STENCILS[st]} n={n3} steps={steps} T={T} block={block} e={ept}", flush=True)
    check_against_oracle(p, rw, ro)


def test_jacobi2d_config_reference_cpu_path(otlib):
    """Jacobi-2D 5-point, 1024x1024, 64 steps, U[0,1) interior from seed, fixed 0
    boundary, T=4, 16x16 block — against the oracle's restatement of the
    reference CPU loop (jacobi-2d.cpp:821-840), in both geometries."""
    for geometry in ("tile", "reference"):
        p = otlib.make_params(otlib.J2D, n=1024, steps=64, T=4, x=16, y=16, e=1, geometry=geometry)
        d = otlib.derive(p)
        init = otlib.fill_reference_uniform(d, 123456)
        rw = [init[otlib.interior_slices(d)].copy()]
        metrics, gpu, ref = check_against_oracle(p, rw)
        print(geometry, d.real_size[0], metrics[0], flush=True)
        if geometry == "reference":
            assert d.real_size[0] == 1020  # gcd(10, 10) truncation kept


def test_graph_replay_equals_eager(otlib):
    p = _params(otlib.J3D, (48, 40, 36), 10, 4, (16, 8, 8), 2)
    d = otlib.derive(p)
    rw, _ = random_inputs(otlib.J3D, domain_shape(d), seed=7)
    eager, _ = run_gpu(p, rw)
    for _ in range(3):
        graph, _ = run_gpu(p, rw, use_graph=True)
        assert np.array_equal(eager[0].view(np.uint32), graph[0].view(np.uint32))


@pytest.mark.parametrize("st", [otlib.J2D, otlib.J3D, otlib.FDTD2D, otlib.RICIAN2D, otlib.TV2D])
def test_time_tile_invariance_bitwise(otlib, st):
    """Overlapped tiling recomputes the same operations in the same order, so any
    (T, tile, ept) must give the T=1 result bit for bit."""
    dims = otlib.stencil_info(st).dims
    n3 = (45, 39, 1) if dims == 2 else (34, 29, 31)
    base = None
    for T, block, ept in [(1, (16, 16, 8), 1), (2, (8, 16, 4), 2), (5, (16, 8, 8), 1), (7, (16, 8, 8), 4)]:
        p = _params(st, n3, 12, T, block, ept)
        d = otlib.derive(p)
        rw, ro = random_inputs(st, domain_shape(d), seed=3)
        out, _ = run_gpu(p, rw, ro)
        if base is None:
            base = out
        else:
            for a, b in zip(base, out):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (T, block, ept)


def test_zero_steps_returns_input(otlib):
    p = _params(otlib.J2D, (33, 17, 1), 0, 2, (16, 16, 1), 1)
    d = otlib.derive(p)
    rw, _ = random_inputs(otlib.J2D, domain_shape(d))
    out, res = run_gpu(p, rw)
    assert res.timing.launches == 0
    assert np.array_equal(out[0], rw[0])


def test_fdtd_physical_sign(otlib):
    p = _params(otlib.FDTD2D, (70, 50, 1), 9, 3, (16, 16, 1), 1, variant=otlib.VARIANT_FDTD_PHYSICAL)
    d = otlib.derive(p)
    rw, _ = random_inputs(otlib.FDTD2D, domain_shape(d), seed=11)
    check_against_oracle(p, rw)


def test_tv_and_rician_parameters(otlib):
    p = _params(otlib.TV2D, (50, 40, 1), 6, 3, (16, 16, 1), 1, coef=[0.1, 0, 0, 0])
    rw, ro = random_inputs(otlib.TV2D, (1, 40, 50), seed=5)
    check_against_oracle(p, rw, ro)
    p = _params(otlib.RICIAN2D, (50, 40, 1), 6, 3, (16, 16, 1), 1, coef=[0.003, 0.1, 2.0, 0])
    rw, ro = random_inputs(otlib.RICIAN2D, (1, 40, 50), seed=6)
    check_against_oracle(p, rw, ro)


@pytest.mark.parametrize("st", [otlib.J1D3, otlib.J1D5, otlib.J2D, otlib.P2D, otlib.J3D, otlib.G3D])
def test_reference_geometry_layout(otlib, st):
    """Reference geometry: padded layout with a ring of mult*T, realSize
    truncated by gcd/lcm, glibc srand(123456) init; the engine computes the full
    realSize domain (the reference's CPU loop) and leaves the ring untouched."""
    dims = otlib.stencil_info(st).dims
    n = {1: 300, 2: 70, 3: 30}[dims]
    block = {1: (32, 1, 1), 2: (16, 12, 1), 3: (8, 8, 8)}[dims]
    p = otlib.make_params(st, n=n, steps=6, T=2, x=block[0], y=block[1], z=block[2], e=1,
                          geometry="reference")
    d = otlib.derive(p)
    init = otlib.fill_reference_uniform(d, 123456)
    res = otlib.run_host(p, [init])
    out = res.fields[0]
    ring = np.ones(out.shape, bool)
    ring[otlib.interior_slices(d)] = False
    assert np.array_equal(out[ring], init[ring])
    ref = oracle.run(st, [init[otlib.interior_slices(d)].copy()], steps=6)[0]
    m = oracle.compare(ref, out[otlib.interior_slices(d)])
    assert m["max_rel"] <= MAX_REL and m["bitwise_equal"], m


def test_enqueue_only_solves_match_synchronous(otlib):
    """OT_RUN_ASYNC enqueues the same loop (eager or graph replay) without
    waiting; back-to-back solves from the same input give the synchronous
    result."""
    p = _params(otlib.J3D, (70, 37, 35), 7, 2, (60, 12, 8), 1)
    d = otlib.derive(p)
    rw, _ = random_inputs(otlib.J3D, domain_shape(d), seed=9)
    ref, _ = run_gpu(p, rw)
    plan = otlib.Plan(p)
    full = otlib.empty_field(d)
    full[otlib.interior_slices(d)] = rw[0]
    for use_graph in (False, True):
        A = otlib.DeviceBuffer.from_array(full)
        B = otlib.DeviceBuffer.from_array(full)
        tm = plan.run([A], [B], use_graph=use_graph, enqueue_only=True)
        otlib.check(otlib.lib().ot_dev_sync())
        out = (B if tm.final_buffer else A).download(full.shape)[otlib.interior_slices(d)]
        assert np.array_equal(out.view(np.uint32), ref[0].view(np.uint32)), use_graph
        A.free()
        B.free()
    plan.close()


FULL_SIZE = [
    # (stencil, extents, steps, tiled (T, block, ept), T=1 (block, ept)): the bench workloads
    (otlib.J3D, (256, 256, 256), 32, (3, (58, 6, 19), 1), ((64, 4, 32), 4)),
    # the deepest time tiles of the 256^3 sweep: the wave kernel at T=5 and the
    # LDS kernel at T=8, down to its smallest tile 8 x 8 x 4
    (otlib.J3D, (256, 256, 256), 32, (5, (54, 2, 20), 1), ((64, 4, 32), 4)),
    # the sweep's best on some boxes: T=4 with 16 footprint rows, 1920 waves
    # (issue balancing on; the T=5 tile above, 8320 waves, runs without it)
    (otlib.J3D, (256, 256, 256), 32, (4, (56, 8, 22), 1), ((64, 4, 32), 4)),
    (otlib.J3D, (256, 256, 256), 32, (8, (16, 8, 16), 1), ((64, 4, 32), 4)),
    (otlib.J3D, (256, 256, 256), 32, (8, (8, 8, 4), 1), ((64, 4, 32), 4)),
    (otlib.G3D, (256, 256, 256), 16, (3, (58, 6, 20), 1), ((64, 4, 32), 4)),
    (otlib.P3D7, (256, 256, 256), 32, (2, (60, 4, 29), 1), ((64, 4, 32), 4)),
    (otlib.FDTD2D, (2048, 2048, 1), 100, (4, (64, 8, 1), 2), ((64, 4, 1), 4)),
    (otlib.J2D, (1024, 1024, 1), 64, (4, (16, 16, 1), 1), ((64, 4, 1), 4)),
    # the 2D wave kernel at the bench's tiles (mask-free path on interior tiles)
    (otlib.FDTD2D, (2048, 2048, 1), 100, (7, (50, 18, 1), 1), ((64, 4, 1), 4)),
    (otlib.FDTD2D, (2048, 2048, 1), 100, (7, (50, 10, 1), 1), ((64, 4, 1), 4)),
    (otlib.J2D, (1024, 1024, 1), 64, (8, (48, 16, 1), 1), ((64, 4, 1), 4)),
    (otlib.J2D, (1024, 1024, 1), 64, (8, (48, 8, 1), 1), ((64, 4, 1), 4)),
    # cs-3d plans only wave tiles, so its T=1 comparison is the bench's T=1 tile
    (otlib.CS3D, (256, 256, 256), 8, (2, (60, 8, 16), 1), ((62, 6, 10), 1)),
]


@pytest.mark.parametrize("case", FULL_SIZE, ids=[f"{otlib.STENCILS[c[0]]}-full-T{c[3][0]}-{c[3][1][0]}x{c[3][1][1]}x{c[3][1][2]}"
                                             for c in FULL_SIZE])
def test_full_size_time_tiling_is_bitwise_t1(otlib, case):
    """At the measured workload sizes the time-tiled solve equals the T=1 solve
    bit for bit (size-independent property: same operations in the same order),
    and the T=1 solve's first step equals the oracle's."""
    st, n3, steps, (T, blk, ept), (blk1, ept1) = case
    p = _params(st, n3, steps, T, blk, ept)
    p1 = _params(st, n3, steps, 1, blk1, ept1)
    d = otlib.derive(p)
    rw, ro = random_inputs(st, domain_shape(d), seed=31)
    print(f"full-size {otlib.STENCILS[st]} {n3} steps={steps} T={T} {blk}", flush=True)
    tiled, _ = run_gpu(p, rw, ro)
    base, _ = run_gpu(p1, rw, ro)
    for a, b in zip(tiled, base):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    one = _params(st, n3, 1, T, blk, ept)
    check_against_oracle(one, rw, ro)


PHASE_CASES = [
    # one case per kernel family (1D LDS, 2D LDS, 2D wave-per-tile, 2D row-streaming,
    # 3D LDS, 3D wave-pipelined with and without its steady-state blocks)
    (otlib.J1D3, (3000, 1, 1), 6, 2, (256, 1, 1), 2),
    (otlib.J2D, (130, 77, 1), 6, 2, (16, 8, 1), 2),
    (otlib.FDTD2D, (130, 77, 1), 6, 2, (60, 12, 1), 1),
    (otlib.J2D, (200, 150, 1), 6, 2, (60, 64, 1), 1),
    (otlib.J3D, (40, 37, 35), 6, 2, (16, 8, 8), 2),
    (otlib.J3D, (70, 20, 60), 6, 3, (58, 6, 16), 1),
    (otlib.P3D27, (70, 25, 29), 6, 2, (60, 8, 8), 1),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", PHASE_CASES, ids=_ids(PHASE_CASES))
@pytest.mark.parametrize("phase", [1, 2, 3])
def test_phase_limited_launches_store_nothing(otlib, case, phase):
    """The timing split of the driver's -p option (1: set-up only, 2: loads and the
    first step, 3: every step but the first, from a synthetic tile) runs each
    kernel family part way and stores no result: every field is left bit for bit
    as it was, and the run completes without an error."""
    st, n3, steps, T, blk, ept = case
    p = _params(st, n3, steps, T, blk, ept, p=phase)
    d = otlib.derive(p)
    rw, ro = random_inputs(st, domain_shape(d), seed=5)
    print(f"phase {phase} {otlib.STENCILS[st]} {n3} T={T} {blk} e{ept}", flush=True)
    out, _ = run_gpu(p, rw, ro)
    for a, b in zip(out, rw):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
