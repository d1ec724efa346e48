"""Pins of the FP64 cost-model oracle (CPU): backpropagation against central
finite differences, the forward pass and the first Adam step against closed
forms, and the normalisation rule."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _random_params(rng, scale=0.3):
    return rng.uniform(-scale, scale, M.NP)


def test_parameter_count_matches_layout():
    assert M.NP == 28 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 10241


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = _random_params(rng)
    X = rng.normal(size=(7, M.F))
    y = rng.normal(size=7)
    mean, istd = rng.normal(size=M.F) * 0.1, rng.uniform(0.5, 2.0, M.F)
    _, g = M.loss_grad(p, X, y, mean, istd)
    # every bias and output weight plus a random sample of the matrices
    idx = list(range(M.NP - 65, M.NP)) + list(rng.choice(M.NP - 65, 300, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NP)
        e[i] = h
        fd = (M.loss(p + e, X, y, mean, istd) - M.loss(p - e, X, y, mean, istd)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_closed_form_single_path():
    # one active path x0 -> h1[0] -> h2[0] -> h3[0] -> out: out = relu(x0) * 2 + 0.5
    P = {k: np.zeros(s) for k, s in M.SHAPES}
    P["W1"][0, 0] = 1.0
    P["W2"][0, 0] = 1.0
    P["W3"][0, 0] = 1.0
    P["W4"][0] = 2.0
    P["b4"][0] = 0.5
    X = np.zeros((4, M.F))
    X[:, 0] = [-1.0, 0.0, 0.25, 3.0]
    out, _ = M.forward(M.pack(P), X)
    assert np.array_equal(out, np.array([0.5, 0.5, 1.0, 6.5]))
    # normalisation shifts and scales the input before the first layer
    out2, _ = M.forward(M.pack(P), X, mean=np.full(M.F, 1.0), inv_std=np.full(M.F, 2.0))
    assert np.array_equal(out2, 2.0 * np.maximum((X[:, 0] - 1.0) * 2.0, 0) + 0.5)


def test_loss_is_mean_squared_error():
    P = {k: np.zeros(s) for k, s in M.SHAPES}
    P["b4"][0] = 1.0
    X = np.zeros((3, M.F))
    y = np.array([1.0, 2.0, 4.0])
    assert M.loss(M.pack(P), X, y) == pytest.approx((0 + 1 + 9) / 3)


def test_first_adam_step_moves_by_lr_times_sign():
    g = np.array([3.0, -0.5, 1e-3, 0.0])
    p = np.zeros(4)
    p1, m, v = M.adam_step(p, g, np.zeros(4), np.zeros(4), 1, lr=0.01)
    # bias-corrected first moments are g and g^2: step = lr * g / (|g| + eps)
    assert np.allclose(p1, -0.01 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_normalization_constant_column_is_centred_only():
    X = np.array([[1.0, 5.0], [3.0, 5.0]])
    X = np.concatenate([X, np.zeros((2, M.F - 2))], 1)
    mu, istd = M.normalization(X)
    assert mu[0] == 2.0 and istd[0] == 1.0  # population sd of {1, 3} is 1
    assert mu[1] == 5.0 and istd[1] == 1.0  # constant column
