"""Pins the FP64 MLP checker of the schedule selector by central differences."""
import numpy as np

from oracle import mlp


def _params(rng, F=7, H=5, C=4):
    return {"W1": rng.standard_normal((F, H)), "b1": rng.standard_normal(H) * 0.1,
            "W2": rng.standard_normal((H, H)), "b2": rng.standard_normal(H) * 0.1,
            "W3": rng.standard_normal((H, C)), "b3": rng.standard_normal(C) * 0.1}


def test_mlp_gradients_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.standard_normal((11, 7))
    q = mlp.softmax(rng.standard_normal((11, 4)))
    _, g = mlp.loss_and_grads(p, x, q)
    eps = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for i in range(0, flat.size, max(1, flat.size // 6)):
            old = flat[i]
            flat[i] = old + eps
            lp, _ = mlp.loss_and_grads(p, x, q)
            flat[i] = old - eps
            lm, _ = mlp.loss_and_grads(p, x, q)
            flat[i] = old
            assert abs((lp - lm) / (2 * eps) - g[k].reshape(-1)[i]) < 1e-6


def test_mlp_hard_labels_are_one_hot_targets():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x = rng.standard_normal((6, 7))
    y = np.array([0, 3, 1, 1, 2, 0])
    q = np.eye(4)[y]
    loss, _ = mlp.loss_and_grads(p, x, q)
    z, _ = mlp.forward(p, x)
    expect = -np.mean(np.log(mlp.softmax(z)[np.arange(6), y]))
    assert abs(loss - expect) < 1e-12


def test_adam_matches_closed_form_first_step():
    p = {"w": np.array([1.0, -2.0])}
    g = {"w": np.array([0.5, -0.25])}
    m, v = {"w": np.zeros(2)}, {"w": np.zeros(2)}
    mlp.adam(p, g, m, v, step=1, lr=0.1)
    # first Adam step moves every coordinate by lr·sign(g) (bias-corrected m/sqrt(v) = g/|g|)
    assert np.allclose(p["w"], [0.9, -1.9], atol=1e-7)
