"""Tile-configuration tuning over the engine's C ABI (include/ot/ot_tune.h).

The reference tunes by expanding a YAML sweep and running one driver process
per configuration (scripts/utils.py:36-244, scripts/benchrunner/benchmarks.py:7-30),
then ranks configurations with an analytic model (scripts/run-model.py) or a
WEKA IBk regressor (scripts/run-weka.py:74-149) and reports how far the
model's pick is from the measured best (run-model.py:575-623,
scripts/parse-weka-results.py:28-35). This module exposes the same steps:

  space(p)              configurations ot_tune_space accepts for p's problem
  features(p)           ot_features vector (the learned model's input)
  analytic_seconds(...) ot_model_analytic (run-model.py restated, MI355X constants)
  measure(...)          ot_tune_measure (median EventElapsed of k solves)
  exhaustive(...)       ot_tune_exhaustive (resumable CSV with the driver keys)
  CostModel             the MLP (HIP training / inference kernels)
  read_csv(path)        rows of a results file as dicts
"""
import csv
import ctypes as C
import math

import numpy as np

import otlib
from otlib import Params, check, lib

FEATURES = otlib.FEATURES


def space(p: Params):
    n = C.c_int32(0)
    check(lib().ot_tune_space(C.byref(p), None, 0, C.byref(n)))
    arr = (Params * max(n.value, 1))()
    check(lib().ot_tune_space(C.byref(p), arr, n.value, C.byref(n)))
    return [arr[i] for i in range(n.value)]


def features(p: Params) -> np.ndarray:
    f = np.zeros(FEATURES, np.float32)
    check(lib().ot_features(C.byref(p), f.ctypes.data_as(otlib._FP)))
    return f


def default_arch() -> otlib.Arch:
    a = otlib.Arch()
    check(lib().ot_arch_default(C.byref(a)))
    return a


def analytic_seconds(p: Params, registers: int, arch: otlib.Arch = None) -> float:
    a = arch if arch is not None else default_arch()
    s = C.c_double(0.0)
    check(lib().ot_model_analytic(C.byref(p), C.byref(a), int(registers), C.byref(s)))
    return s.value


def measure(p: Params, rw_a, rw_b, ro=(), warmup=1, reps=3, time_limit_s=0.0):
    """(median seconds, registers, timed_out) of one solve on DeviceBuffers."""
    t = C.c_double(0.0)
    regs = C.c_int32(0)
    to = C.c_int32(0)
    check(lib().ot_tune_measure(C.byref(p), otlib._ptr_array([b.ptr for b in rw_a]),
                                otlib._ptr_array([b.ptr for b in rw_b]), otlib._ptr_array([b.ptr for b in ro]),
                                warmup, reps, time_limit_s, C.byref(t), C.byref(regs), C.byref(to)))
    return t.value, regs.value, bool(to.value)


def exhaustive(p: Params, csv_path: str, warmup=1, reps=3, time_limit_s=0.0):
    """(best Params, best seconds, configurations measured by this call)."""
    best = Params()
    bs = C.c_double(0.0)
    m = C.c_int32(0)
    check(lib().ot_tune_exhaustive(C.byref(p), csv_path.encode(), warmup, reps, time_limit_s, C.byref(best),
                                   C.byref(bs), C.byref(m)))
    return best, bs.value, m.value


def read_csv(path: str):
    """Rows of a results file (plain or gzip-compressed) as dicts."""
    if path.endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            return list(csv.DictReader(f))
    with open(path) as f:
        return list(csv.DictReader(f))


def params_of_row(row) -> Params:
    """Params of one results-file row (tile geometry extents)."""
    return otlib.make_params(row["Stencil"], nx=int(row["Nx"]), ny=int(row["Ny"]), nz=int(row["Nz"]),
                             problem_size=int(row["Nx"]), steps=int(row["Time Steps"]),
                             geometry=int(row["Geometry"]), T=int(row["Time Tile Size"]),
                             x=int(row["Block Size X"]), y=int(row["Block Size Y"]),
                             z=int(row["Block Size Z"]), e=int(row["Elements/Thread"]))


def label_of_row(row) -> float:
    """log(seconds per grid point per time step) of one measured row."""
    pts = int(row["Nx"]) * int(row["Ny"]) * int(row["Nz"])
    return math.log(float(row["EventElapsed"]) / (pts * int(row["Time Steps"])))


class CostModel:
    """The learned cost model: OT_FEATURES -> 3 x 64 ReLU -> log(s per point-step)."""

    def __init__(self, seed: int = 1, handle=None):
        self.h = C.c_void_p(handle) if handle is not None else C.c_void_p()
        if handle is None:
            check(lib().ot_model_create(seed, C.byref(self.h)))

    @classmethod
    def load(cls, path: str) -> "CostModel":
        h = C.c_void_p()
        check(lib().ot_model_load(path.encode(), C.byref(h)))
        return cls(handle=h.value)

    def save(self, path: str) -> None:
        check(lib().ot_model_save(self.h, path.encode()))

    @property
    def param_count(self) -> int:
        n = C.c_int32(0)
        check(lib().ot_model_param_count(self.h, C.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        check(lib().ot_model_get_params(self.h, p.ctypes.data_as(otlib._FP), p.size))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(lib().ot_model_set_params(self.h, p.ctypes.data_as(otlib._FP), p.size))

    def set_normalization(self, mean: np.ndarray, inv_std: np.ndarray) -> None:
        mean = np.ascontiguousarray(mean, np.float32)
        inv_std = np.ascontiguousarray(inv_std, np.float32)
        check(lib().ot_model_set_normalization(self.h, mean.ctypes.data_as(otlib._FP),
                                               inv_std.ctypes.data_as(otlib._FP)))

    def fit_normalization(self, X: np.ndarray) -> None:
        X = np.ascontiguousarray(X, np.float32)
        check(lib().ot_model_fit_normalization(self.h, X.ctypes.data_as(otlib._FP), X.shape[0]))

    def loss_grad(self, X: np.ndarray, y: np.ndarray, want_grad=True):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.param_count, np.float32) if want_grad else None
        check(lib().ot_model_loss_grad(self.h, X.ctypes.data_as(otlib._FP), y.ctypes.data_as(otlib._FP),
                                       X.shape[0], loss.ctypes.data_as(otlib._FP),
                                       g.ctypes.data_as(otlib._FP) if g is not None else None))
        return float(loss[0]), g

    def train(self, X, y, epochs=100, batch=256, lr=1e-3, seed=1) -> float:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().ot_model_train(self.h, X.ctypes.data_as(otlib._FP), y.ctypes.data_as(otlib._FP), X.shape[0],
                                   epochs, batch, lr, seed, loss.ctypes.data_as(otlib._FP)))
        return float(loss[0])

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        check(lib().ot_model_predict(self.h, X.ctypes.data_as(otlib._FP), X.shape[0], y.ctypes.data_as(otlib._FP)))
        return y

    def predict_params(self, p: Params) -> float:
        v = np.zeros(1, np.float32)
        check(lib().ot_predict(self.h, C.byref(p), v.ctypes.data_as(otlib._FP)))
        return float(v[0])

    def close(self):
        if self.h:
            lib().ot_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
