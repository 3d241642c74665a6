"""mlp_oracle - FP64 CPU restatement of the learned proposal's network.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network's forward pass, gradient and Adam step; never by the engine.

The network has no counterpart in the reference (whose conditional tests run
in a fixed column order, kernel3.cu:70-109); it scores gene pairs for the
learned test window.  12 inputs -> 64 (ReLU) -> 64 (ReLU) -> 1 (sigmoid),
binary cross-entropy against soft labels y in [0, 1], mean over the batch.

Parameter vector layout (5057 floats):
  W1 [64][12] at 0, b1 [64] at 768, W2 [64][64] at 832, b2 [64] at 4928,
  w3 [64] at 4992, b3 at 5056.

Pins (tests/test_mlp_oracle.py): the gradient against central finite
differences of the loss in FP64; the closed form for zero weights (output
sigmoid(b3), dL/db3 = mean(sigmoid(b3) - y), every other gradient 0); the
first Adam step, which moves every parameter by lr * g / (|g| + eps).
"""
import numpy as np

NIN, NH = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAMS = OFF_B3 + 1


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[OFF_W1:OFF_B1].reshape(NH, NIN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(NH, NH),
            P[OFF_B2:OFF_W3], P[OFF_W3:OFF_B3], P[OFF_B3])


def logits(P, X):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def forward(P, X):
    return 1.0 / (1.0 + np.exp(-logits(P, X)))


def loss(P, X, Y):
    z = logits(P, X)
    Y = np.asarray(Y, np.float64)
    return float(np.mean(np.logaddexp(0.0, z) - Y * z))


def gradient(P, X, Y):
    """(dL/dP, L) of the mean binary cross-entropy."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    n = len(X)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    dz = (1.0 / (1.0 + np.exp(-z)) - Y) / n
    d2 = np.outer(dz, w3) * (a2 > 0)
    d1 = (d2 @ W2) * (a1 > 0)
    g = np.zeros(NPARAMS)
    g[OFF_W1:OFF_B1] = (d1.T @ X).reshape(-1)
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).reshape(-1)
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dz
    g[OFF_B3] = dz.sum()
    return g, float(np.mean(np.logaddexp(0.0, z) - Y * z))


def adam_step(P, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1); returns (P, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v
