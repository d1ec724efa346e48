"""FP64 CPU restatement of the learned cost model (test infrastructure only).

Checker for the HIP kernels in ics-12-overlapped-tiling_amd/csrc/mlp.hip: the
same 28 -> 64 -> 64 -> 64 -> 1 ReLU perceptron over the same flat parameter
layout (W1[28x64] b1 W2[64x64] b2 W3[64x64] b3 W4[64] b4, row-major, input
index first), the same feature normalisation x' = (x - mean) * inv_std, the
mean-squared-error loss, its gradient by backpropagation, and one Adam step
(beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected). Everything is float64.

The reference has no counterpart (its learned ranking is WEKA IBk in a JVM,
scripts/run-weka.py:74-149, not vendored). The gradient is pinned by central
finite differences and the forward pass by closed forms in
tests/test_mlp_oracle.py.
"""
import numpy as np

F = 28
H = 64
SHAPES = [("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H,)), ("b4", (1,))]
NP = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    assert flat.size == NP, flat.size
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def forward(flat, X, mean=None, inv_std=None):
    """Returns (prediction [n], cache for backward)."""
    P = unpack(flat)
    X = np.asarray(X, np.float64)
    mean = np.zeros(F) if mean is None else np.asarray(mean, np.float64)
    inv_std = np.ones(F) if inv_std is None else np.asarray(inv_std, np.float64)
    x0 = (X - mean) * inv_std
    z1 = x0 @ P["W1"] + P["b1"]
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ P["W2"] + P["b2"]
    a2 = np.maximum(z2, 0.0)
    z3 = a2 @ P["W3"] + P["b3"]
    a3 = np.maximum(z3, 0.0)
    out = a3 @ P["W4"] + P["b4"][0]
    return out, (x0, z1, a1, z2, a2, z3, a3)


def loss(flat, X, y, mean=None, inv_std=None):
    out, _ = forward(flat, X, mean, inv_std)
    return float(np.mean((out - np.asarray(y, np.float64)) ** 2))


def loss_grad(flat, X, y, mean=None, inv_std=None):
    """(mean squared error, gradient with respect to the flat parameters)."""
    P = unpack(flat)
    y = np.asarray(y, np.float64)
    out, (x0, z1, a1, z2, a2, z3, a3) = forward(flat, X, mean, inv_std)
    n = out.shape[0]
    d_out = 2.0 * (out - y) / n
    g = {}
    g["W4"] = a3.T @ d_out
    g["b4"] = np.array([d_out.sum()])
    dz3 = np.outer(d_out, P["W4"]) * (z3 > 0)
    g["W3"] = a2.T @ dz3
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ P["W3"].T) * (z2 > 0)
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ P["W2"].T) * (z1 > 0)
    g["W1"] = x0.T @ dz1
    g["b1"] = dz1.sum(0)
    return float(np.mean((out - y) ** 2)), pack(g)


def adam_step(flat, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update at step t (1-based). Returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mhat = m / (1.0 - b1 ** t)
    vhat = v / (1.0 - b2 ** t)
    return flat - lr * mhat / (np.sqrt(vhat) + eps), m, v


def normalization(X):
    """Per-column mean and inverse population standard deviation (1 for a
    constant column), as ot_model_fit_normalization computes them."""
    X = np.asarray(X, np.float64)
    mu = X.mean(0)
    sd = np.sqrt(((X - mu) ** 2).mean(0))
    return mu, np.where(sd > 1e-6, 1.0 / np.where(sd > 1e-6, sd, 1.0), 1.0)
