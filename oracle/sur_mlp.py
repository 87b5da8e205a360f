"""FP64 numpy restatement of the delayed-acceptance surrogate's network.

TEST INFRASTRUCTURE ONLY (the checker of the device training kernels; never
imported by the product).  The network of one sublattice with F descriptor
features is

    g(x) = w3 . tanh(W2 tanh(W1 x + b1) + b2) + v . x        (hidden width 64)

with the flat FP32 parameter vector [W1 (64 x F, row-major), b1, W2 (64 x 64,
row-major), b2, w3, v].  For pairs (A_r, B_r) with exact energy differences y_r
the training loss is mean_r ((g(B_r) - g(A_r)) - y_r / scale)^2.
This has no counterpart in the reference (SCEMonteCarlo.jl has no surrogate).
"""
from __future__ import annotations

import numpy as np

H = 64


def unpack(p, F):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((H, F), (H,), (H, H), (H,), (H,), (F,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == len(p), "parameter vector length"
    return out


def forward(p, F, X):
    """g(x) for the rows of X (n x >= F); FP64."""
    W1, b1, W2, b2, w3, v = unpack(p, F)
    X = np.asarray(X, np.float64)[:, :F]
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + X @ v


def loss_and_grad(p, F, A, B, y, scale):
    """Loss and its gradient with respect to the flat parameter vector (FP64)."""
    W1, b1, W2, b2, w3, v = unpack(p, F)
    n = len(y)
    X = np.concatenate([np.asarray(A, np.float64)[:, :F], np.asarray(B, np.float64)[:, :F]])
    Z1 = X @ W1.T + b1
    H1 = np.tanh(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = np.tanh(Z2)
    g = H2 @ w3 + X @ v
    res = (g[n:] - g[:n]) - np.asarray(y, np.float64) / scale
    loss = float(np.mean(res * res))
    dg = np.concatenate([-2.0 * res / n, 2.0 * res / n])
    gw3 = H2.T @ dg
    gv = X.T @ dg
    dZ2 = np.outer(dg, w3) * (1.0 - H2 * H2)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (1.0 - H1 * H1)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, gv])
    return loss, grad


def glorot(F, rng):
    a1, a2 = np.sqrt(6.0 / (F + H)), np.sqrt(6.0 / (2 * H))
    return np.concatenate([rng.uniform(-a1, a1, H * F), np.zeros(H), rng.uniform(-a2, a2, H * H),
                           np.zeros(H), rng.uniform(-0.1, 0.1, H), np.zeros(F)]).astype(np.float32)
