"""FP64 numpy restatement of the learned-m MLP -- TEST INFRASTRUCTURE ONLY.

No reference counterpart (the network is new in this build).  The layout is
the one of include/hrnn_dev.h (HRNN_MLP_PARAMS): W1[64][32] b1[64] W2[64][64]
b2[64] W3[16][64] b3[16]; ReLU hidden layers, softmax output, mean
cross-entropy loss.  Pinned by central finite differences of the loss
(tests/test_oracle_mlp.py): the analytic gradient must be the derivative.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 32, 64, 16
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_OUT, N_H)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[1])
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    return np.concatenate(parts).astype(np.float32)


def forward(p, X):
    P = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0)
    lg = h2 @ P["W3"].T + P["b3"]
    lg = lg - lg.max(1, keepdims=True)
    e = np.exp(lg)
    prob = e / e.sum(1, keepdims=True)
    return prob, (X, z1, h1, z2, h2)


def loss(p, X, y):
    prob, _ = forward(p, X)
    return float(-np.log(np.maximum(prob[np.arange(len(y)), y], 1e-300)).mean())


def grad(p, X, y):
    P = unpack(p)
    prob, (X, z1, h1, z2, h2) = forward(p, X)
    B = X.shape[0]
    d = prob.copy()
    d[np.arange(B), y] -= 1.0
    d /= B
    gW3 = d.T @ h2
    gb3 = d.sum(0)
    dz2 = (d @ P["W3"]) * (z2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * (z1 > 0)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
