"""FP64 checker for the schedule-selector MLP (TEST INFRASTRUCTURE ONLY).

The selector network has no counterpart in the reference; this numpy
restatement of its maths — standardise, relu(x·W1+b1), relu(·W2+b2), ·W3+b3,
softmax cross-entropy against soft targets, backward, Adam — is what the GPU
kernels (armsparse_amd/csrc/mlp.hip) are checked against, and it is itself
pinned by central finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np


def forward(p: dict, x: np.ndarray):
    x = x.astype(np.float64)
    z1 = x @ p["W1"] + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ p["W3"] + p["b3"]
    return z3, (x, h1, h2)


def softmax(z):
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def loss_and_grads(p: dict, x: np.ndarray, q: np.ndarray):
    """Mean soft-target cross-entropy and its gradients w.r.t. every parameter."""
    z3, (x, h1, h2) = forward(p, x)
    n = x.shape[0]
    pr = softmax(z3)
    loss = float(-(q * np.log(pr + 1e-300)).sum() / n)
    d3 = (pr - q) / n
    g = {"W3": h2.T @ d3, "b3": d3.sum(0)}
    d2 = (d3 @ p["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ p["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ d1, d1.sum(0)
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    for k in p:
        gi = g[k] + wd * p[k]
        m[k] = b1 * m[k] + (1 - b1) * gi
        v[k] = b2 * v[k] + (1 - b2) * gi * gi
        p[k] = p[k] - lr * (m[k] / c1) / (np.sqrt(v[k] / c2) + eps)
