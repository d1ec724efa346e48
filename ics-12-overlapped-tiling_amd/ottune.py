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
  exhaustive_supervised the same sweep in a child process under a hang watchdog
  CostModel             the MLP (HIP training / inference kernels)
  read_csv(path)        rows of a results file as dicts
"""
import csv
import ctypes as C
import json
import math
import os
import signal
import subprocess
import sys
import time

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


CSV_COLUMNS = ["Stencil", "Dimensions", "Nx", "Ny", "Nz", "Time Steps", "Geometry", "Time Tile Size",
               "Block Size X", "Block Size Y", "Block Size Z", "Elements/Thread", "Num Blocks X", "Num Blocks Y",
               "Num Blocks Z", "Threads", "LDS Bytes", "Shared Size", "Register Usage", "Wave Kernel",
               "EventElapsed", "Status"]


def config_key(p: Params):
    """The identity of one configuration in a results file: stencil, extents,
    steps, geometry and tile (the columns ot_tune_exhaustive resumes on)."""
    d = otlib.derive(p)
    return (otlib.STENCILS[p.stencil], d.real_size[0], d.real_size[1], d.real_size[2], p.time_steps, p.geometry,
            p.time_tile, p.block_x, p.block_y, p.block_z, p.elems_per_thread)


def _row_key(r):
    return (r["Stencil"],) + tuple(int(r[c]) for c in ("Nx", "Ny", "Nz", "Time Steps", "Geometry", "Time Tile Size",
                                                       "Block Size X", "Block Size Y", "Block Size Z",
                                                       "Elements/Thread"))


def _data_rows(path: str) -> int:
    try:
        with open(path) as f:
            return max(0, sum(1 for _ in f) - 1)
    except OSError:
        return 0


def _append_status_row(path: str, p: Params, status: str) -> None:
    """Record configuration p with no measurement (EventElapsed -1) and the
    given Status, so that a resumed sweep skips it."""
    d = otlib.derive(p)
    new = not os.path.exists(path) or os.path.getsize(path) == 0
    with open(path, "a") as f:
        if new:
            f.write(",".join(CSV_COLUMNS) + "\n")
        vals = [otlib.STENCILS[p.stencil], d.dims, d.real_size[0], d.real_size[1], d.real_size[2], p.time_steps,
                p.geometry, p.time_tile, p.block_x, p.block_y, p.block_z, p.elems_per_thread, d.grid_blocks[0],
                d.grid_blocks[1], d.grid_blocks[2], d.threads_per_block, d.lds_bytes, d.shared_size, 0,
                int(features(p)[23]), -1, status]
        f.write(",".join(str(v) for v in vals) + "\n")


def _params_json(p: Params) -> str:
    return json.dumps({n: (list(getattr(p, n)) if n == "coef" else getattr(p, n)) for n, _ in Params._fields_})


def _params_from_json(text: str) -> Params:
    p = Params()
    for n, v in json.loads(text).items():
        if n == "coef":
            for i, c in enumerate(v):
                p.coef[i] = c
        else:
            setattr(p, n, v)
    return p


def exhaustive_supervised(p: Params, csv_path: str, warmup=1, reps=3, time_limit_s=0.0, watchdog_s=15.0,
                          startup_s=300.0, max_restarts=50, child_cmd=None, log=None):
    """ot_tune_exhaustive in a child process under a hang watchdog, as the
    reference runs every configuration in its own process and terminates one
    that runs past 15 s (scripts/utils.py:130-145). The sweep writes one row
    per configuration; when the file gains no row for watchdog_s seconds
    (startup_s for the child's first row), the child's whole process group is
    killed, the configuration it was on (the first of the space missing from
    the file) is recorded with Status "hang" ("crash" when the child died by
    itself) and a fresh child resumes the sweep. This process only reads the
    file and enumerates the space on the host, so it never touches the GPU.
    child_cmd replaces the child's command line (a sweep stand-in for tests).
    Returns (best Params or None, best seconds, {"restarts", "hung": [keys]})."""
    sp = space(p)
    keys = [config_key(q) for q in sp]
    cmd = child_cmd or [sys.executable, os.path.abspath(__file__), "sweep-child", _params_json(p), csv_path,
                        str(warmup), str(reps), repr(float(time_limit_s))]
    hung, restarts = [], 0
    while True:
        proc = subprocess.Popen(cmd, start_new_session=True, stdout=subprocess.DEVNULL)
        rows, last = _data_rows(csv_path), time.monotonic()
        limit, killed = startup_s, False
        while proc.poll() is None:
            time.sleep(0.05)
            n = _data_rows(csv_path)
            if n != rows:
                rows, last, limit = n, time.monotonic(), watchdog_s
            elif time.monotonic() - last > limit:
                os.killpg(proc.pid, signal.SIGKILL)
                proc.wait()
                killed = True
                break
        if proc.returncode == 0 and not killed:
            break
        done = {_row_key(r) for r in read_csv(csv_path)} if os.path.exists(csv_path) else set()
        nxt = next((q for q, k in zip(sp, keys) if k not in done), None)
        if nxt is None:
            break
        status = "hang" if killed else "crash"
        _append_status_row(csv_path, nxt, status)
        hung.append(config_key(nxt))
        if log:
            log(f"watchdog: {status} at T={nxt.time_tile} {nxt.block_x}x{nxt.block_y}x{nxt.block_z} "
                f"e{nxt.elems_per_thread}; restarting the sweep")
        restarts += 1
        if restarts > max_restarts:
            raise otlib.OtError(otlib.OT_EHIP, f"sweep restarted {restarts} times without finishing")
    best, best_s = None, float("inf")
    want = set(keys)
    for r in read_csv(csv_path):
        if r["Status"] == "ok" and _row_key(r) in want and float(r["EventElapsed"]) < best_s:
            best_s = float(r["EventElapsed"])
            best = next(q for q, k in zip(sp, keys) if k == _row_key(r))
    return best, best_s, {"restarts": restarts, "hung": hung}


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


if __name__ == "__main__" and len(sys.argv) == 7 and sys.argv[1] == "sweep-child":
    # the child of exhaustive_supervised: one resumable exhaustive sweep
    _p = _params_from_json(sys.argv[2])
    exhaustive(_p, sys.argv[3], warmup=int(sys.argv[4]), reps=int(sys.argv[5]), time_limit_s=float(sys.argv[6]))
