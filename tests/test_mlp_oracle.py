"""Pins of the FP64 network oracle (oracle/mlp_oracle.py) on the CPU."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _net(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, scale, M.NPARAMS)
    X = rng.normal(0, 1, (37, M.NIN))
    Y = rng.uniform(0, 1, 37)
    return P, X, Y


def test_gradient_matches_finite_differences():
    P, X, Y = _net(1)
    g, L = M.gradient(P, X, Y)
    assert abs(L - M.loss(P, X, Y)) < 1e-14
    rng = np.random.default_rng(2)
    # every bias, the last layer and a sample of each weight matrix
    idx = list(range(M.OFF_B1, M.OFF_W2)) + list(range(M.OFF_B2, M.NPARAMS))
    idx += list(rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 60, replace=False))
    idx += list(rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 120, replace=False))
    h = 1e-6
    for q in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += h
        Pm[q] -= h
        fd = (M.loss(Pp, X, Y) - M.loss(Pm, X, Y)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])), (q, fd, g[q])


def test_planted_mistake_is_caught():
    # a gradient with the ReLU mask of layer 1 dropped must fail the same check
    P, X, Y = _net(3)
    g, _ = M.gradient(P, X, Y)
    W1, b1, W2, b2, w3, b3 = M.unpack(P)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    z = np.maximum(a2, 0) @ w3 + b3
    dz = (1 / (1 + np.exp(-z)) - Y) / len(X)
    d1_wrong = (np.outer(dz, w3) * (a2 > 0)) @ W2  # no (a1 > 0)
    bad = d1_wrong.sum(0)
    h = 1e-6
    fd = np.zeros(M.NH)
    for j in range(M.NH):
        Pp, Pm = P.copy(), P.copy()
        Pp[M.OFF_B1 + j] += h
        Pm[M.OFF_B1 + j] -= h
        fd[j] = (M.loss(Pp, X, Y) - M.loss(Pm, X, Y)) / (2 * h)
    assert np.allclose(fd, g[M.OFF_B1:M.OFF_W2], atol=1e-7)
    assert not np.allclose(fd, bad, atol=1e-7)


def test_zero_weights_closed_form():
    P = np.zeros(M.NPARAMS)
    P[M.OFF_B3] = 0.7
    X = np.random.default_rng(4).normal(0, 1, (11, M.NIN))
    Y = np.linspace(0, 1, 11)
    s = 1 / (1 + np.exp(-0.7))
    assert np.allclose(M.forward(P, X), s, rtol=0, atol=1e-15)
    g, L = M.gradient(P, X, Y)
    assert abs(g[M.OFF_B3] - np.mean(s - Y)) < 1e-15
    assert np.count_nonzero(g[:M.OFF_B3]) == 0
    assert abs(L - np.mean(np.log1p(np.exp(0.7)) - 0.7 * Y)) < 1e-14


def test_first_adam_step_is_sign_step():
    rng = np.random.default_rng(5)
    P = rng.normal(0, 1, M.NPARAMS)
    g = rng.normal(0, 1e-3, M.NPARAMS)
    P1, m, v = M.adam_step(P, g, np.zeros_like(P), np.zeros_like(P), 1, 0.01)
    assert np.allclose(P - P1, 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
