"""ctypes wrapper of the CPU oracle (libot_oracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker (or the timed CPU baseline). The oracle restates the
reference's CPU loops (see ot_oracle.c for the per-stencil citations).

Arrays are numpy float32/float64 arrays of the domain (interior only, shape
(nz, ny, nx) with 1 for unused axes); padding with the zero ring of width P
happens here.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libot_oracle.so")

DIMS = {0: 1, 1: 1, 2: 1, 3: 2, 4: 3, 5: 2, 6: 3, 7: 3, 8: 2, 9: 2, 10: 3, 11: 2, 12: 2, 13: 3}
RADIUS = {0: 1, 1: 2, 2: 3}
NUM_RW = {8: 3}
NUM_RO = {7: 1, 11: 1, 12: 2}

_lib = None


def load(path: str):
    """Load an oracle library build (the shipped one, or a copy built from
    edited sources by the tests that check the pins catch planted mistakes)."""
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built: make -C {HERE}")
    L = C.CDLL(path)
    for suffix, ct in (("f32", C.c_float), ("f64", C.c_double)):
        fn = getattr(L, f"oracle_run_{suffix}")
        fn.restype = C.c_int
        fn.argtypes = [C.c_int] * 6 + [C.POINTER(C.POINTER(ct)), C.POINTER(C.POINTER(ct)),
                                      C.POINTER(C.c_float), C.c_int]
    L.oracle_g2d_double_sqrt.restype = None
    L.oracle_g2d_double_sqrt.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_float)]
    L.oracle_reference_init.restype = None
    L.oracle_reference_init.argtypes = [C.c_int] * 6 + [C.c_uint, C.POINTER(C.c_float)]
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load(LIB_PATH)
    return _lib


def pad(interior: np.ndarray, dims: int, P: int) -> np.ndarray:
    """Embed a (nz, ny, nx) interior in the oracle's zero-ring box."""
    widths = [(0, 0)] * 3
    for ax in range(dims):
        widths[2 - ax] = (P, P)
    return np.ascontiguousarray(np.pad(interior, widths))


def unpad(box: np.ndarray, dims: int, P: int) -> np.ndarray:
    sl = [slice(None)] * 3
    for ax in range(dims):
        sl[2 - ax] = slice(P, box.shape[2 - ax] - P)
    return box[tuple(sl)]


def run(stencil: int, rw, ro=(), steps: int = 1, coef=(0, 0, 0, 0), variant: int = 0,
        dtype=np.float32, timed: bool = False, library=None):
    """Advance `steps` time steps. rw/ro: lists of (nz, ny, nx) interior arrays.
    Returns the list of rw interiors after the steps (and the seconds spent in
    the C loop when timed=True). library: an oracle build from load() (default:
    the shipped one)."""
    dims = DIMS[stencil]
    P = max(RADIUS.get(stencil, 1), 1)
    nz, ny, nx = rw[0].shape
    boxes = [pad(np.asarray(a, dtype=dtype), dims, P) for a in rw]
    roboxes = [pad(np.asarray(a, dtype=dtype), dims, P) for a in ro]
    ct = C.c_float if dtype == np.float32 else C.c_double
    L = library if library is not None else lib()
    fn = L.oracle_run_f32 if dtype == np.float32 else L.oracle_run_f64
    rwp = (C.POINTER(ct) * 3)(*[b.ctypes.data_as(C.POINTER(ct)) for b in boxes])
    rop = (C.POINTER(ct) * 2)(*[b.ctypes.data_as(C.POINTER(ct)) for b in roboxes])
    cf = (C.c_float * 4)(*[float(c) for c in (list(coef) + [0, 0, 0, 0])[:4]])
    t0 = time.perf_counter()
    st = fn(stencil, nx, ny, nz, P, steps, rwp, rop, cf, variant)
    dt = time.perf_counter() - t0
    if st != 0:
        raise MemoryError("oracle allocation failed")
    out = [unpad(b, dims, P).copy() for b in boxes]
    return (out, dt) if timed else out


def g2d_double_sqrt(u: np.ndarray, steps: int) -> np.ndarray:
    """Gradient-2D with the reference's double-precision sqrt overload."""
    box = pad(np.asarray(u, dtype=np.float32), 2, 1)
    _, ny, nx = u.shape
    lib().oracle_g2d_double_sqrt(nx, ny, 1, steps, box.ctypes.data_as(C.POINTER(C.c_float)))
    return unpad(box, 2, 1).copy()


def reference_init(dims: int, padded: tuple, P: int, extra_tail: int = 0, seed: int = 123456):
    """glibc srand/rand init over a padded box (sz, sy, sx), ring P."""
    sz, sy, sx = padded
    out = np.zeros(padded, dtype=np.float32)
    lib().oracle_reference_init(dims, sx, sy, sz, P, extra_tail, seed,
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def compare(host: np.ndarray, dev: np.ndarray) -> dict:
    """Error metrics of dev against host: the reference's L2 norms
    (compareResults, jacobi-2d.cpp:517-551) plus max-abs and max-relative."""
    h = np.asarray(host, dtype=np.float64)
    d = np.asarray(dev, dtype=np.float64)
    diff = h - d
    err = float(np.sqrt(np.sum(diff * diff)))
    ref = float(np.sqrt(np.sum(h * h)))
    finite = np.isfinite(h) & np.isfinite(d)
    same_nonfinite = bool(np.all((np.isfinite(h) == np.isfinite(d))))
    absd = np.abs(diff[finite]) if finite.any() else np.zeros(1)
    denom = np.maximum(np.abs(h[finite]), 1e-30) if finite.any() else np.ones(1)
    rel = absd / denom
    # max relative error over points whose magnitude is not negligible
    scale = np.max(np.abs(h[finite])) if finite.any() else 0.0
    mask = np.abs(h[finite]) > 1e-6 * max(scale, 1e-30) if finite.any() else np.zeros(1, bool)
    return {
        "error_norm": err,
        "ref_norm": ref,
        "l2_rel": err / ref if ref > 0 else float("inf"),
        "max_abs": float(absd.max()) if absd.size else 0.0,
        "max_rel": float(rel[mask].max()) if mask.any() else 0.0,
        "bitwise_equal": bool(np.array_equal(np.asarray(host, np.float32).view(np.uint32),
                                             np.asarray(dev, np.float32).view(np.uint32))),
        "same_nonfinite": same_nonfinite,
        "correctness": "PASSED" if (ref >= 1e-7 and err / ref <= 1e-2) else "FAILED",
    }
