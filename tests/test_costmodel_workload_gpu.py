"""The learned cost model on its own workload (MI355X): the shipped network
(tuning/mi355x/mlp.bin, 3 x 64 ReLU, trained on the GPU by the engine's own
kernels) ranks every configuration of each held-out shape of the committed
autotuner timings (tuning/mi355x/timings.csv.gz: exhaustive sweeps of 140
random shapes over all 14 stencils; every fifth shape of each stencil is held
out of training) with the HIP inference kernels. Each pick's slowdown is its
measured time over the shape's exhaustive best, the metric of the reference's
model evaluation (scripts/run-model.py:575-623,
scripts/parse-weka-results.py:28-35).

The method without the learned part is the analytic model (run-model.py
restated); the test shows what the network gains over it.
"""
import os

import numpy as np
import pytest

import costmodel
import ottune

pytestmark = pytest.mark.gpu

TUNING = os.path.join(os.path.dirname(os.path.abspath(costmodel.__file__)), "tuning", "mi355x")


def test_learned_picks_on_held_out_shapes(otlib):
    groups = costmodel.load(TUNING)
    _, held_out = costmodel.split(groups)
    assert len(held_out) == 28
    m = ottune.CostModel.load(os.path.join(TUNING, "mlp.bin"))
    arch = ottune.default_arch()
    sweeps = costmodel._sweeps(TUNING)
    recs = [costmodel.rank_shape(m, k, groups[k], arch, sweeps) for k in held_out]
    m.close()
    s = costmodel.summarize(recs)
    top1, top3 = s["learned_top1_slowdown"], s["learned_top3_slowdown"]
    ana, ana1pct = s["analytic_top1_slowdown"], s["analytic_top1pct_slowdown"]
    ratio = s["learned_top3_tuning_ratio"]
    print(f"held-out shapes {len(recs)}: learned best-of-3 mean {top3['mean']:.3f} "
          f"(within 10% on {top3['within_10pct']:.0%}), learned top-1 mean {top1['mean']:.3f}; "
          f"analytic top-1 mean {ana['mean']:.3f}, analytic best of top 1% {ana1pct['mean']:.3f}; "
          f"predict + measure top 3 costs {ratio['mean']:.4f} of the exhaustive sweep's solve time", flush=True)
    # the learned pick (predict everything, measure the top 3) lands within 5%
    # of the exhaustive best on average, at under 1/50 of its tuning time
    assert top3["mean"] <= 1.05
    assert top3["within_10pct"] >= 0.85
    assert ratio["mean"] <= 1 / 50
    # what the network gains over the analytic model it replaces
    assert top1["mean"] < ana["mean"] and top3["mean"] < ana1pct["mean"]
    gain = np.array([r["analytic_top1_slowdown"] / r["learned_top1_slowdown"] for r in recs])
    assert float(np.exp(np.mean(np.log(gain)))) > 2.0
