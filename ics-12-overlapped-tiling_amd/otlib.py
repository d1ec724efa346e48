"""ctypes binding of the engine's C ABI (include/ot/ot.h).

Mirrors the reference's generator parameter surface (GeneratorParams,
opencl/jacobi-2d/jacobi-2d.cpp:21-127) and its host time loop
(opencl/jacobi-2d/jacobi-2d.cpp:888-948). The library is libot.so built in
this directory; loading fails loudly when it is missing — there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libot.so")

STENCILS = [
    "jacobi-1d", "jacobi-1d-5pt", "jacobi-1d-7pt", "jacobi-2d", "jacobi-3d", "poisson-2d",
    "poisson-3d", "poisson-3d-rhs", "fdtd-2d", "gradient-2d", "gradient-3d",
    "rician-denoise-2d", "cdsc-tv-update-2d", "cs-3d",
]
(J1D3, J1D5, J1D7, J2D, J3D, P2D, P3D27, P3D7, FDTD2D, G2D, G3D, RICIAN2D, TV2D, CS3D) = range(14)
GEOM_TILE, GEOM_REFERENCE = 0, 1
VARIANT_FDTD_PHYSICAL = 1
VARIANT_TIME_ONLY_GMEM, VARIANT_TIME_ONLY_INSTR, VARIANT_TIME_ONLY_SMEM, VARIANT_DUMP_CLOCKS = 2, 4, 8, 16
VARIANT_NO_BALANCE = 32

OT_OK, OT_EINVAL, OT_ERESOURCE, OT_EHIP, OT_EUNSUPPORTED = range(5)


class OtError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"ot status {status}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("stencil", C.c_int32), ("problem_size", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("time_steps", C.c_int32), ("time_tile", C.c_int32),
        ("block_x", C.c_int32), ("block_y", C.c_int32), ("block_z", C.c_int32),
        ("elems_per_thread", C.c_int32), ("geometry", C.c_int32), ("phase_limit", C.c_int32),
        ("variant", C.c_int32), ("coef", C.c_float * 4),
    ]


class Derived(C.Structure):
    _fields_ = [
        ("dims", C.c_int32), ("radius", C.c_int32), ("padding", C.c_int32), ("halo", C.c_int32),
        ("comps_per_block", C.c_int32 * 3), ("real_per_block", C.c_int32 * 3),
        ("size_lcm", C.c_int32), ("real_size", C.c_int32 * 3), ("padded_size", C.c_int32 * 3),
        ("num_blocks", C.c_int32 * 3), ("tile", C.c_int32 * 3), ("grid_blocks", C.c_int32 * 3),
        ("shared_size", C.c_int32), ("lds_bytes", C.c_int32), ("threads_per_block", C.c_int32),
        ("launches", C.c_int32), ("num_rw", C.c_int32), ("num_ro", C.c_int32),
        ("field_elems", C.c_int64), ("stride_y", C.c_int64), ("stride_z", C.c_int64),
        ("origin", C.c_int64),
        ("total_fp_per_block", C.c_double), ("useful_fp_per_block", C.c_double),
        ("useful_ratio", C.c_double), ("total_fp_steps", C.c_double),
        ("useful_fp_steps", C.c_double),
        ("global_loads_per_block", C.c_int64), ("global_stores_per_block", C.c_int64),
        ("shared_loads_per_block", C.c_int64), ("shared_stores_per_block", C.c_int64),
        ("registers", C.c_int32), ("reserved0", C.c_int32),
    ]


class Timing(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("event_s", C.c_double), ("kernel_s", C.c_double),
                ("kernel_full_s", C.c_double), ("launches", C.c_int32),
                ("full_launches", C.c_int32), ("final_buffer", C.c_int32),
                ("reserved", C.c_int32)]


RUN_GRAPH, RUN_TIME_LAUNCHES, RUN_ASYNC = 1, 2, 8


class Image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pitch", C.c_int32),
                ("reserved", C.c_int32), ("offset", C.c_int64)]


class StencilInfo(C.Structure):
    _fields_ = [("stencil", C.c_int32), ("dims", C.c_int32), ("radius", C.c_int32),
                ("num_rw", C.c_int32), ("num_ro", C.c_int32), ("flops_per_point", C.c_int32),
                ("loads_per_point", C.c_int32), ("ref_padding_mult", C.c_int32),
                ("name", C.c_char_p)]


class Arch(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "total_lds", "max_waves_per_cu", "max_threads_per_cu", "max_blocks_per_cu", "max_regs_per_cu",
        "num_cu", "clock_hz", "c_load", "c_op", "b_gmem", "b_smem", "l_gmem", "l_smem", "launch_penalty",
        "wave_size", "max_regs_per_thread", "l1_bytes", "sfu_cycles")]


_FP = C.POINTER(C.c_float)
_PP = C.POINTER(_FP)
FEATURES = 28

# (name, restype, argtypes) of every entry point declared in include/ot/*.h
SIGNATURES = [
    ("ot_version", C.c_char_p, []),
    ("ot_last_error", C.c_char_p, []),
    ("ot_device_query", C.c_int, [C.POINTER(C.c_int32)] * 3 + [C.c_char_p, C.c_int32]),
    ("ot_stencil_query", C.c_int, [C.c_int32, C.POINTER(StencilInfo)]),
    ("ot_stencil_from_name", C.c_int, [C.c_char_p, C.POINTER(C.c_int32)]),
    ("ot_params_default", C.c_int, [C.c_int32, C.POINTER(Params)]),
    ("ot_derive", C.c_int, [C.POINTER(Params), C.POINTER(Derived)]),
    ("ot_plan_create", C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("ot_plan_derived", C.c_int, [C.c_void_p, C.POINTER(Derived)]),
    ("ot_plan_destroy", None, [C.c_void_p]),
    ("ot_plan_stream", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("ot_run", C.c_int, [C.c_void_p, _PP, _PP, _PP, C.c_void_p, C.c_int32, C.POINTER(Timing)]),
    ("ot_launch", C.c_int, [C.c_void_p, _PP, _PP, _PP, C.c_int32, C.c_void_p]),
    ("ot_plan_clocks", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    ("ot_naive_run", C.c_int, [C.POINTER(Params), _PP, _PP, _PP, C.c_void_p, C.POINTER(Timing)]),
    ("ot_batch_layout", C.c_int, [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(Image),
                                  C.POINTER(C.c_int64)]),
    ("ot_batch_create", C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(Image), C.POINTER(C.c_void_p)]),
    ("ot_batch_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("ot_batch_run", C.c_int, [C.c_void_p, _PP, _PP, _PP, C.c_void_p, C.c_int32, C.POINTER(Timing)]),
    ("ot_batch_destroy", None, [C.c_void_p]),
    ("ot_batch_run_static", C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(Image), _PP, _PP, _PP, C.c_void_p,
                                      C.POINTER(Timing)]),
    ("ot_fill_rician_phantom", C.c_int, [_FP, C.c_int32, C.POINTER(Image), C.c_uint64, C.c_float]),
    ("ot_dev_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("ot_dev_free", C.c_int, [C.c_void_p]),
    ("ot_memcpy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("ot_memcpy_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("ot_memcpy_d2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("ot_dev_memset", C.c_int, [C.c_void_p, C.c_int32, C.c_size_t]),
    ("ot_dev_sync", C.c_int, []),
    ("ot_fill_reference_uniform", C.c_int, [_FP, C.POINTER(Derived), C.c_uint32]),
    ("ot_fill_uniform", C.c_int, [_FP, C.POINTER(Derived), C.c_uint64, C.c_float, C.c_float]),
    ("ot_fill_sinusoid_noise", C.c_int, [_FP, C.POINTER(Derived), C.c_uint64, C.c_float]),
    ("ot_fill_gaussian_pulse", C.c_int, [_FP, C.POINTER(Derived), C.c_float, C.c_float]),
    ("ot_selftest_rsq", C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    # ot_tune.h
    ("ot_tune_space", C.c_int, [C.POINTER(Params), C.POINTER(Params), C.c_int32, C.POINTER(C.c_int32)]),
    ("ot_tune_measure", C.c_int, [C.POINTER(Params), _PP, _PP, _PP, C.c_int32, C.c_int32, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("ot_tune_exhaustive", C.c_int, [C.POINTER(Params), C.c_char_p, C.c_int32, C.c_int32, C.c_double,
                                     C.POINTER(Params), C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    ("ot_arch_default", C.c_int, [C.POINTER(Arch)]),
    ("ot_model_analytic", C.c_int, [C.POINTER(Params), C.POINTER(Arch), C.c_int32, C.POINTER(C.c_double)]),
    ("ot_features", C.c_int, [C.POINTER(Params), _FP]),
    ("ot_model_create", C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    ("ot_model_destroy", None, [C.c_void_p]),
    ("ot_model_param_count", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("ot_model_get_params", C.c_int, [C.c_void_p, _FP, C.c_int32]),
    ("ot_model_set_params", C.c_int, [C.c_void_p, _FP, C.c_int32]),
    ("ot_model_set_normalization", C.c_int, [C.c_void_p, _FP, _FP]),
    ("ot_model_fit_normalization", C.c_int, [C.c_void_p, _FP, C.c_int32]),
    ("ot_model_loss_grad", C.c_int, [C.c_void_p, _FP, _FP, C.c_int32, _FP, _FP]),
    ("ot_model_train", C.c_int, [C.c_void_p, _FP, _FP, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                 _FP]),
    ("ot_model_predict", C.c_int, [C.c_void_p, _FP, C.c_int32, _FP]),
    ("ot_predict", C.c_int, [C.c_void_p, C.POINTER(Params), _FP]),
    ("ot_model_save", C.c_int, [C.c_void_p, C.c_char_p]),
    ("ot_model_load", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
]

_lib = None


def lib():
    """Load libot.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run __graft_entry__.build() or make -C "
                               f"{HERE}")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OT_OK:
        raise OtError(status, lib().ot_last_error().decode())


def stencil_id(name_or_id) -> int:
    if isinstance(name_or_id, int):
        return name_or_id
    return STENCILS.index(name_or_id)


def stencil_info(stencil) -> StencilInfo:
    info = StencilInfo()
    check(lib().ot_stencil_query(stencil_id(stencil), C.byref(info)))
    return info


def make_params(stencil, **kw) -> Params:
    p = Params()
    check(lib().ot_params_default(stencil_id(stencil), C.byref(p)))
    aliases = {"n": "problem_size", "t": "time_steps", "steps": "time_steps", "T": "time_tile",
               "s": "time_tile", "x": "block_x", "y": "block_y", "z": "block_z",
               "e": "elems_per_thread", "ept": "elems_per_thread", "p": "phase_limit"}
    for k, v in kw.items():
        k = aliases.get(k, k)
        if k == "coef":
            for i, c in enumerate(v):
                p.coef[i] = c
        elif k == "geometry" and isinstance(v, str):
            p.geometry = GEOM_REFERENCE if v == "reference" else GEOM_TILE
        else:
            setattr(p, k, int(v))
    return p


def derive(p: Params) -> Derived:
    d = Derived()
    check(lib().ot_derive(C.byref(p), C.byref(d)))
    return d


def layout_shape(d: Derived) -> tuple:
    """numpy shape (z, y, x) of one padded field buffer."""
    return (d.padded_size[2], d.padded_size[1], d.stride_y)


def interior_slices(d: Derived) -> tuple:
    """(z, y, x) slices selecting the domain inside a padded field buffer."""
    o = d.origin
    oz = o // d.stride_z if d.dims > 2 else 0
    o -= oz * d.stride_z
    oy = o // d.stride_y if d.dims > 1 else 0
    ox = o - oy * d.stride_y
    return (slice(oz, oz + d.real_size[2]), slice(oy, oy + d.real_size[1]),
            slice(ox, ox + d.real_size[0]))


def empty_field(d: Derived) -> np.ndarray:
    return np.zeros(layout_shape(d), dtype=np.float32)


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_FP)


def fill_reference_uniform(d: Derived, seed: int = 123456) -> np.ndarray:
    a = empty_field(d)
    check(lib().ot_fill_reference_uniform(_fp(a), C.byref(d), seed))
    return a


def fill_uniform(d: Derived, seed: int, lo: float, hi: float) -> np.ndarray:
    a = empty_field(d)
    check(lib().ot_fill_uniform(_fp(a), C.byref(d), seed, lo, hi))
    return a


def fill_sinusoid_noise(d: Derived, seed: int, amp: float) -> np.ndarray:
    a = empty_field(d)
    check(lib().ot_fill_sinusoid_noise(_fp(a), C.byref(d), seed, amp))
    return a


def fill_gaussian_pulse(d: Derived, amplitude: float, sigma: float) -> np.ndarray:
    a = empty_field(d)
    check(lib().ot_fill_gaussian_pulse(_fp(a), C.byref(d), amplitude, sigma))
    return a


class DeviceBuffer:
    """Device allocation through the engine's HIP helpers."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().ot_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        check(lib().ot_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        check(lib().ot_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer") -> None:
        assert other.nbytes == self.nbytes
        check(lib().ot_memcpy_d2d(self.ptr, other.ptr, self.nbytes))

    def free(self) -> None:
        if self.ptr:
            check(lib().ot_dev_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.ot_dev_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _ptr_array(ptrs):
    arr = (_FP * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = C.cast(C.c_void_p(p), _FP)
    return arr


class Plan:
    """An ot_plan: derived geometry plus the compiled kernels it launches."""

    def __init__(self, params: Params):
        self.params = params
        h = C.c_void_p()
        check(lib().ot_plan_create(C.byref(params), C.byref(h)))
        self.handle = h
        self.derived = Derived()
        check(lib().ot_plan_derived(h, C.byref(self.derived)))

    def run(self, rw_a, rw_b, ro=(), use_graph=False, time_launches=False, stream=None,
            enqueue_only=False) -> Timing:
        """rw_a/rw_b/ro: sequences of DeviceBuffer. Returns the Timing.
        enqueue_only: OT_RUN_ASYNC (no wait, no timing; sync with ot_dev_sync)."""
        a = _ptr_array([b.ptr for b in rw_a])
        b = _ptr_array([x.ptr for x in rw_b])
        r = _ptr_array([x.ptr for x in ro]) if ro else None
        tm = Timing()
        mode = ((RUN_GRAPH if use_graph else 0) | (RUN_TIME_LAUNCHES if time_launches else 0) |
                (RUN_ASYNC if enqueue_only else 0))
        check(lib().ot_run(self.handle, a, b, r, stream, mode, C.byref(tm)))
        return tm

    def stream(self):
        """The plan's own stream, on which graph replays with stream=None run."""
        h = C.c_void_p()
        check(lib().ot_plan_stream(self.handle, C.byref(h)))
        return h.value

    def launch(self, src, dst, ro=(), base_time=0, stream=None) -> None:
        s = _ptr_array([x.ptr for x in src])
        d = _ptr_array([x.ptr for x in dst])
        r = _ptr_array([x.ptr for x in ro]) if ro else None
        check(lib().ot_launch(self.handle, s, d, r, base_time, stream))

    def close(self):
        if self.handle:
            lib().ot_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StreamTimer:
    """A HIP event pair on one stream (None = the null stream): the time the GPU
    spends on the work enqueued on that stream between start() and stop().
    Events come from the HIP runtime libot.so itself links."""

    _hip = None

    def __init__(self, stream=None):
        if StreamTimer._hip is None:
            lib()
            h = C.CDLL("libamdhip64.so.7")
            h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
            h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
            h.hipEventSynchronize.argtypes = [C.c_void_p]
            h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
            h.hipEventDestroy.argtypes = [C.c_void_p]
            StreamTimer._hip = h
        self.stream = stream
        self.ev = [C.c_void_p(), C.c_void_p()]
        for e in self.ev:
            if StreamTimer._hip.hipEventCreate(C.byref(e)) != 0:
                raise OtError(OT_EHIP, "hipEventCreate failed")

    def start(self):
        if StreamTimer._hip.hipEventRecord(self.ev[0], self.stream) != 0:
            raise OtError(OT_EHIP, "hipEventRecord failed")

    def stop(self) -> float:
        """Record the end event, wait for it and return the seconds between the two."""
        h = StreamTimer._hip
        if h.hipEventRecord(self.ev[1], self.stream) != 0 or h.hipEventSynchronize(self.ev[1]) != 0:
            raise OtError(OT_EHIP, "hipEventRecord/hipEventSynchronize failed")
        ms = C.c_float()
        if h.hipEventElapsedTime(C.byref(ms), self.ev[0], self.ev[1]) != 0:
            raise OtError(OT_EHIP, "hipEventElapsedTime failed")
        return ms.value * 1e-3

    def close(self):
        for e in self.ev:
            if e:
                StreamTimer._hip.hipEventDestroy(e)
        self.ev = [C.c_void_p(), C.c_void_p()]


@dataclass
class RunResult:
    fields: list          # host arrays (padded layout) of the rw fields after all steps
    timing: Timing
    derived: Derived


def run_host(params: Params, rw_init, ro_init=(), use_graph=False, repeat=1) -> RunResult:
    """Upload host fields, run the whole time loop on the GPU, download the result.
    With repeat > 1 the loop is run again from the same initial state and the
    last timing is returned (results of every repetition are identical)."""
    plan = Plan(params)
    d = plan.derived
    shape = layout_shape(d)
    a = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in rw_init]
    b = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in rw_init]
    ro = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in ro_init]
    tm = None
    for _ in range(repeat):
        for buf, x in zip(a, rw_init):
            buf.upload(np.ascontiguousarray(x, dtype=np.float32))
        tm = plan.run(a, b, ro, use_graph=use_graph)
    src = a if tm.final_buffer == 0 else b
    out = [buf.download(shape) for buf in src]
    for buf in a + b + ro:
        buf.free()
    plan.close()
    return RunResult(out, tm, d)


def run_naive_host(params: Params, rw_init, ro_init=()) -> RunResult:
    """The same problem through the non-tiled one-step-per-launch kernels."""
    d = derive(params)
    shape = layout_shape(d)
    a = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in rw_init]
    b = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in rw_init]
    ro = [DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32)) for x in ro_init]
    pa = _ptr_array([x.ptr for x in a])
    pb = _ptr_array([x.ptr for x in b])
    pr = _ptr_array([x.ptr for x in ro]) if ro else None
    tm = Timing()
    check(lib().ot_naive_run(C.byref(params), pa, pb, pr, None, C.byref(tm)))
    src = a if tm.final_buffer == 0 else b
    out = [buf.download(shape) for buf in src]
    for buf in a + b + ro:
        buf.free()
    return RunResult(out, tm, d)
