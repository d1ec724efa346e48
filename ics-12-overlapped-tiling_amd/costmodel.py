"""Learned tile-configuration cost model: data collection, training, evaluation.

    python costmodel.py collect  --out DIR [--per-stencil K] [--seed S]
    python costmodel.py train    --data DIR --model FILE
    python costmodel.py evaluate --data DIR --model FILE --report FILE
    python costmodel.py loso     --data DIR --report FILE

collect  runs the exhaustive autotuner (ot_tune_exhaustive, in a child process
         under a hang watchdog) over seeded random problem shapes of every
         stencil (1D up to 2^24 points, 2D up to 4096^2, 3D up to 384^3) into
         DIR/timings.csv (resumable: rerunning continues where it stopped) and
         records the wall time of every shape's sweep in DIR/sweeps.json.
train    fits the MLP (ot_model_*: HIP forward / backward / Adam kernels) on
         the shapes not held out (80 %, held out per stencil by shape).
evaluate ranks every held-out shape's configuration space with the model and
         with the analytic model (ot_model_analytic, the restatement of
         scripts/run-model.py), and reports each pick's slowdown against the
         measured best (the metric of run-model.py:575-623 and
         scripts/parse-weka-results.py:28-35) and the tuning time of
         "predict, then measure the top k" against the exhaustive sweep.
loso     the reference's leave-one-stencil-out protocol: one model per
         stencil, trained without it, ranking all of its shapes.

The reference's learned precedent is WEKA IBk over 17 features with a
leave-one-stencil-out protocol (scripts/extract-features.py:40-209,
scripts/run-weka.py:74-149); this model is a 28-feature MLP (one-hot stencil,
shape, tile, stencil descriptors and the analytic model's estimate as a prior)
and its held-out set is unseen shapes of every stencil; `loso` repeats the
reference's protocol with the one-hot block zeroed.
"""
import argparse
import json
import math
import os
import random
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import otlib  # noqa: E402
import ottune  # noqa: E402

STEPS = 24
HOLDOUT_EVERY = 5  # shape i of a stencil is held out when i % 5 == 4 (20 %)


def shapes(per_stencil: int, seed: int):
    """[(stencil name, (nx, ny, nz))] drawn log-uniformly per dimensionality."""
    rng = random.Random(seed)
    out = []
    for st, name in enumerate(otlib.STENCILS):
        dims = otlib.stencil_info(st).dims
        for _ in range(per_stencil):
            if dims == 1:
                n3 = (int(2 ** rng.uniform(16, 24)), 1, 1)
            elif dims == 2:
                n3 = (int(2 ** rng.uniform(8, 12)), int(2 ** rng.uniform(8, 12)), 1)
            else:
                n3 = tuple(int(2 ** rng.uniform(6, math.log2(384))) for _ in range(3))
            out.append((name, n3))
    return out


def shape_key(name, n3):
    return f"{name}:{n3[0]}x{n3[1]}x{n3[2]}"


def base_params(name, n3):
    # the sweep's layout plan: a T=1 tile every stencil accepts (cs-3d plans
    # only wave-pipelined tiles)
    x, y = (62, 6) if name == "cs-3d" else (16, 4)
    return otlib.make_params(name, nx=n3[0], ny=n3[1], nz=n3[2], problem_size=n3[0], steps=STEPS, T=1, x=x,
                             y=y, z=4, e=1)


def collect(out_dir: str, per_stencil: int, seed: int, time_limit: float, stencils=None, watchdog=15.0):
    """Exhaustive sweeps of the shapes drawn for `seed`, appended to
    out_dir/timings.csv. `stencils` (names) restricts the sweep to those
    stencils' shapes; the draw itself is unchanged, so a restricted run
    measures exactly the shapes a full run would for them."""
    os.makedirs(out_dir, exist_ok=True)
    csv_path = os.path.join(out_dir, "timings.csv")
    sweeps_path = os.path.join(out_dir, "sweeps.json")
    sweeps = json.load(open(sweeps_path)) if os.path.exists(sweeps_path) else {}
    todo = [(name, n3) for name, n3 in shapes(per_stencil, seed) if not stencils or name in stencils]
    for i, (name, n3) in enumerate(todo):
        key = shape_key(name, n3)
        t0 = time.perf_counter()
        before = ottune._data_rows(csv_path)
        # each shape's sweep runs in a child process under the hang watchdog
        best, best_s, info = ottune.exhaustive_supervised(base_params(name, n3), csv_path, warmup=1, reps=3,
                                                          time_limit_s=time_limit, watchdog_s=watchdog,
                                                          log=lambda m: print(m, flush=True))
        measured = ottune._data_rows(csv_path) - before - len(info["hung"])
        wall = time.perf_counter() - t0
        if best is None:
            print(f"[{i + 1}/{len(todo)}] {key}: no configuration measured", flush=True)
            continue
        if measured:
            sweeps[key] = {"wall_s": wall, "configs": measured}
            json.dump(sweeps, open(sweeps_path, "w"), indent=1)
        print(f"[{i + 1}/{len(todo)}] {key}: {measured} configs in {wall:.1f} s, best {best_s * 1e3:.3f} ms "
              f"(T={best.time_tile} {best.block_x}x{best.block_y}x{best.block_z} e{best.elems_per_thread})",
              flush=True)


def load(data_dir: str):
    """Rows grouped by shape: {key: [row, ...]} of measured ('ok') rows."""
    path = os.path.join(data_dir, "timings.csv")
    rows = ottune.read_csv(path if os.path.exists(path) else path + ".gz")
    groups = {}
    for r in rows:
        if r["Status"] != "ok":
            continue
        key = shape_key(r["Stencil"], (int(r["Nx"]), int(r["Ny"]), int(r["Nz"])))
        groups.setdefault(key, []).append(r)
    return groups


def split(groups):
    """(train keys, held-out keys): every fifth shape of each stencil, in
    sorted order, is held out."""
    by_stencil = {}
    for k in sorted(groups):
        by_stencil.setdefault(k.split(":")[0], []).append(k)
    train, test = [], []
    for keys in by_stencil.values():
        for i, k in enumerate(keys):
            (test if i % HOLDOUT_EVERY == HOLDOUT_EVERY - 1 else train).append(k)
    return train, test


ONEHOT = 13  # leading one-hot stencil columns of the feature vector


def matrix(groups, keys, mask_onehot=False):
    X, y = [], []
    for k in keys:
        for r in groups[k]:
            X.append(ottune.features(ottune.params_of_row(r)))
            y.append(ottune.label_of_row(r))
    X = np.asarray(X, np.float32)
    if mask_onehot:
        X[:, :ONEHOT] = 0.0
    return X, np.asarray(y, np.float32)


def train(data_dir: str, model_path: str, epochs: int, batch: int, lr: float, seed: int):
    groups = load(data_dir)
    tr, te = split(groups)
    X, y = matrix(groups, tr)
    m = ottune.CostModel(seed=seed)
    m.fit_normalization(X)
    l0, _ = m.loss_grad(X, y, want_grad=False)
    t0 = time.perf_counter()
    loss = m.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    dt = time.perf_counter() - t0
    Xt, yt = matrix(groups, te)
    lt, _ = m.loss_grad(Xt, yt, want_grad=False) if len(yt) else (float("nan"), None)
    m.save(model_path)
    info = {"train_rows": int(len(y)), "test_rows": int(len(yt)), "train_shapes": len(tr), "test_shapes": len(te),
            "initial_mse": l0, "train_mse": loss, "test_mse": lt, "epochs": epochs, "batch": batch, "lr": lr,
            "train_seconds": dt}
    print(json.dumps(info), flush=True)
    return info


def solve_cost(r, warmup=1, reps=3):
    """Seconds of GPU solves the sweep spent on one configuration."""
    return (warmup + reps) * float(r["EventElapsed"])


def rank_shape(m, key, rows, arch, sweeps, top_k=(1, 3), mask_onehot=False):
    """Learned and analytic picks for one shape's measured rows."""
    params = [ottune.params_of_row(r) for r in rows]
    t = np.array([float(r["EventElapsed"]) for r in rows])
    best = float(t.min())
    t0 = time.perf_counter()
    X = np.stack([ottune.features(p) for p in params])
    if mask_onehot:
        X[:, :ONEHOT] = 0.0
    pred = m.predict(X)
    infer_s = time.perf_counter() - t0
    order = np.argsort(pred, kind="stable")
    ana = np.array([ottune.analytic_seconds(p, int(r["Register Usage"]), arch) for p, r in zip(params, rows)])
    aorder = np.argsort(ana, kind="stable")
    n_top1pct = max(1, int(math.ceil(0.01 * len(rows))))
    exhaustive_s = float(sum(solve_cost(r) for r in rows))
    rec = {"shape": key, "configs": len(rows), "best_s": best,
           "exhaustive_solve_s": exhaustive_s, "exhaustive_wall_s": sweeps.get(key, {}).get("wall_s"),
           "inference_s": infer_s,
           "analytic_top1_slowdown": float(t[aorder[0]] / best),
           "analytic_top1pct_slowdown": float(t[aorder[:n_top1pct]].min() / best),
           "best_config": {n: rows[int(t.argmin())][n] for n in
                           ("Time Tile Size", "Block Size X", "Block Size Y", "Block Size Z", "Elements/Thread")}}
    for kk in top_k:
        sel = order[:kk]
        rec[f"learned_top{kk}_slowdown"] = float(t[sel].min() / best)
        learned_cost = infer_s + float(sum(solve_cost(rows[i]) for i in sel))
        rec[f"learned_top{kk}_tuning_ratio"] = learned_cost / exhaustive_s
        if rec["exhaustive_wall_s"]:
            rec[f"learned_top{kk}_wall_ratio"] = learned_cost / rec["exhaustive_wall_s"]
    return rec


def summarize(per_shape, top_k=(1, 3)):
    def agg(field):
        v = np.array([r[field] for r in per_shape if r.get(field) is not None], float)
        return {"mean": float(v.mean()), "median": float(np.median(v)), "max": float(v.max()),
                "within_10pct": float(np.mean(v <= 1.10))} if len(v) else None

    return {f: agg(f) for f in ["analytic_top1_slowdown", "analytic_top1pct_slowdown"] +
            [f"learned_top{k}_slowdown" for k in top_k] + [f"learned_top{k}_tuning_ratio" for k in top_k] +
            [f"learned_top{k}_wall_ratio" for k in top_k]}


def _sweeps(data_dir):
    path = os.path.join(data_dir, "sweeps.json")
    return json.load(open(path)) if os.path.exists(path) else {}


def evaluate(data_dir: str, model_path: str, report_path: str):
    groups = load(data_dir)
    sweeps = _sweeps(data_dir)
    _, te = split(groups)
    m = ottune.CostModel.load(model_path)
    arch = ottune.default_arch()
    per_shape = [rank_shape(m, k, groups[k], arch, sweeps) for k in te]
    summary = summarize(per_shape)
    report = {"held_out_shapes": len(per_shape), "summary": summary, "per_shape": per_shape}
    json.dump(report, open(report_path, "w"), indent=1)
    print(json.dumps(summary), flush=True)
    return report


def loso(data_dir: str, report_path: str, epochs: int, batch: int, lr: float, seed: int):
    """Leave-one-stencil-out (the reference's protocol, scripts/run-weka.py:74-149):
    for every stencil, train on all shapes of the other stencils and rank every
    shape of the left-out one."""
    groups = load(data_dir)
    sweeps = _sweeps(data_dir)
    arch = ottune.default_arch()
    stencils = sorted({k.split(":")[0] for k in groups})
    per_stencil, all_shapes = {}, []
    for st in stencils:
        tr = [k for k in groups if k.split(":")[0] != st]
        te = sorted(k for k in groups if k.split(":")[0] == st)
        X, y = matrix(groups, tr, mask_onehot=True)
        m = ottune.CostModel(seed=seed)
        m.fit_normalization(X)
        m.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
        recs = [rank_shape(m, k, groups[k], arch, sweeps, mask_onehot=True) for k in te]
        m.close()
        per_stencil[st] = summarize(recs)
        all_shapes += recs
        s1 = per_stencil[st]["learned_top1_slowdown"]
        s3 = per_stencil[st]["learned_top3_slowdown"]
        sa = per_stencil[st]["analytic_top1_slowdown"]
        print(f"{st}: learned top-1 mean {s1['mean']:.3f}, top-3 mean {s3['mean']:.3f}, "
              f"analytic top-1 mean {sa['mean']:.3f} over {len(recs)} shapes", flush=True)
    report = {"protocol": "leave-one-stencil-out", "epochs": epochs, "summary": summarize(all_shapes),
              "per_stencil": per_stencil, "per_shape": all_shapes}
    json.dump(report, open(report_path, "w"), indent=1)
    print(json.dumps(report["summary"]), flush=True)
    return report


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--out", required=True)
    c.add_argument("--per-stencil", type=int, default=5)
    c.add_argument("--seed", type=int, default=2024)
    c.add_argument("--time-limit", type=float, default=0.05, help="watchdog: seconds per solve")
    c.add_argument("--stencils", default="", help="comma-separated stencil names to sweep (default: all)")
    c.add_argument("--watchdog", type=float, default=15.0,
                   help="seconds without a finished configuration before the sweep's process is restarted")
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--model", required=True)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    e = sub.add_parser("evaluate")
    e.add_argument("--data", required=True)
    e.add_argument("--model", required=True)
    e.add_argument("--report", required=True)
    lo = sub.add_parser("loso")
    lo.add_argument("--data", required=True)
    lo.add_argument("--report", required=True)
    lo.add_argument("--epochs", type=int, default=150)
    lo.add_argument("--batch", type=int, default=256)
    lo.add_argument("--lr", type=float, default=1e-3)
    lo.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    if a.cmd == "collect":
        collect(a.out, a.per_stencil, a.seed, a.time_limit, [x for x in a.stencils.split(",") if x], a.watchdog)
    elif a.cmd == "train":
        train(a.data, a.model, a.epochs, a.batch, a.lr, a.seed)
    elif a.cmd == "evaluate":
        evaluate(a.data, a.model, a.report)
    else:
        loso(a.data, a.report, a.epochs, a.batch, a.lr, a.seed)


if __name__ == "__main__":
    main()
