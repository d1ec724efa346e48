"""The cost model's HIP kernels (csrc/mlp.hip) against the FP64 oracle
(oracle/mlp_oracle.py): forward, loss, gradient, one Adam step, normalisation,
training progress, save/load and the params -> features -> prediction path.
FP32 kernels against FP64 references: tolerances are stated per check."""
import os
import sys

import numpy as np
import pytest

import otlib
import ottune

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, M.F)).astype(np.float32)
    y = (0.5 * X[:, 0] - 0.25 * X[:, 3] * X[:, 5] + np.sin(X[:, 7])).astype(np.float32)
    return X, y


def test_forward_loss_grad_match_oracle(otlib):
    X, y = _data(300, 1)
    m = ottune.CostModel(seed=11)
    p = m.get_params()
    assert p.size == M.NP
    mean = np.linspace(-0.2, 0.2, M.F).astype(np.float32)
    istd = np.linspace(0.5, 1.5, M.F).astype(np.float32)
    m.set_normalization(mean, istd)
    pred = m.predict(X)
    ref, _ = M.forward(p, X, mean, istd)
    # FP32 accumulation over <= 64 terms per layer: 1e-5 relative to the output scale
    assert np.max(np.abs(pred - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))
    loss, g = m.loss_grad(X, y)
    rloss, rg = M.loss_grad(p, X, y, mean, istd)
    assert loss == pytest.approx(rloss, rel=1e-5)
    # gradient: within 1e-4 of the largest component (FP32 sums over 300 rows)
    assert np.max(np.abs(g - rg)) <= 1e-4 * np.max(np.abs(rg))


def test_loss_grad_chunked_over_4096_rows(otlib):
    X, y = _data(9000, 2)  # three chunks of the whole-set pass
    m = ottune.CostModel(seed=5)
    p = m.get_params()
    loss, g = m.loss_grad(X, y)
    rloss, rg = M.loss_grad(p, X, y)
    assert loss == pytest.approx(rloss, rel=1e-4)
    assert np.max(np.abs(g - rg)) <= 2e-4 * np.max(np.abs(rg))


def test_one_full_batch_adam_step_matches_oracle(otlib):
    X, y = _data(256, 3)
    m = ottune.CostModel(seed=7)
    p0 = m.get_params().astype(np.float64)
    m.train(X, y, epochs=1, batch=256, lr=1e-3, seed=1)
    p1 = m.get_params()
    _, rg = M.loss_grad(p0, X, y)
    rp1, _, _ = M.adam_step(p0, rg, np.zeros(M.NP), np.zeros(M.NP), 1, lr=1e-3)
    # the first step is lr * g / (|g| + eps): compare where the sign of g is
    # well determined in FP32 (|g| above 1e-3 of the largest component)
    sel = np.abs(rg) > 1e-3 * np.max(np.abs(rg))
    assert sel.sum() > 1000
    assert np.max(np.abs(p1[sel] - rp1[sel])) <= 1e-6
    assert np.max(np.abs(p1 - rp1)) <= 2.0e-3 + 1e-6  # elsewhere at most one opposite step


def test_training_reduces_loss_and_is_reproducible(otlib):
    X, y = _data(2048, 4)
    losses = []
    for _ in range(2):
        m = ottune.CostModel(seed=3)
        m.fit_normalization(X)
        l0, _ = m.loss_grad(X, y, want_grad=False)
        l1 = m.train(X, y, epochs=40, batch=128, lr=2e-3, seed=9)
        losses.append((l0, l1, m.get_params()))
    (l0, l1, pa), (_, l1b, pb) = losses
    assert l1 < 0.2 * l0
    assert l1 == l1b and np.array_equal(pa, pb)  # fixed-order reductions, seeded shuffle


def test_fit_normalization_matches_oracle(otlib):
    X, _ = _data(500, 5)
    X[:, 2] = 4.0  # constant column
    m = ottune.CostModel(seed=1)
    m.fit_normalization(X)
    mu, istd = M.normalization(X)
    p = m.get_params()
    pred = m.predict(X)
    ref, _ = M.forward(p, X, mu, istd)
    assert np.max(np.abs(pred - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))


def test_save_load_and_params_path(otlib, tmp_path):
    X, y = _data(512, 6)
    m = ottune.CostModel(seed=2)
    m.fit_normalization(X)
    m.train(X, y, epochs=3, batch=64, lr=1e-3, seed=2)
    path = str(tmp_path / "m.bin")
    m.save(path)
    m2 = ottune.CostModel.load(path)
    assert np.array_equal(m.predict(X), m2.predict(X))
    p = otlib.make_params("jacobi-3d", nx=256, ny=256, nz=256, problem_size=256, steps=32, T=3, x=58, y=6, z=20,
                          e=1)
    v = m2.predict_params(p)
    assert v == m2.predict(ottune.features(p)[None, :])[0]
    with pytest.raises(otlib.OtError):
        open(str(tmp_path / "bad.bin"), "wb").write(b"nope")
        ottune.CostModel.load(str(tmp_path / "bad.bin"))
