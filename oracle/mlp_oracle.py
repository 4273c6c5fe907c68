"""Float64 CPU oracle of the learned K-th-score model (3-layer MLP, in -> 64 -> 64 -> 1, ReLU).
TEST INFRASTRUCTURE ONLY: imported by tests/ to check the GPU forward pass and gradients
(bm25mi_mlp_predict / bm25mi_mlp_gradient); never by the product package.

The model has no counterpart in the reference. Parameter layout (flat, as the library's):
W1[64][in] b1[64] W2[64][64] b2[64] W3[64] b3[1]. Loss = (1/n) sum (y - target)^2.
Pinned by central finite differences in float64 (tests/test_learned.py).
"""

import numpy as np

H = 64


def unpack(params, n_in):
    p = np.asarray(params, np.float64)
    o = 0
    W1 = p[o:o + H * n_in].reshape(H, n_in); o += H * n_in
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(params, x):
    x = np.asarray(x, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(params, x.shape[1])
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3 + b3


def loss(params, x, target):
    d = forward(params, x) - np.asarray(target, np.float64)
    return float(np.mean(d * d))


def gradient(params, x, target):
    """Backpropagation of the mean squared error; returns (grad in the flat layout, loss)."""
    x = np.asarray(x, np.float64)
    n, n_in = x.shape
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    d = y - np.asarray(target, np.float64)
    dy = 2.0 * d / n
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, W3) * (z2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return grad, float(np.mean(d * d))
