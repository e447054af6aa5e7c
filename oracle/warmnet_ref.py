"""CPU reference (float64 numpy) of the warm-start network's forward pass, loss
and gradient, the checker for the GPU kernels behind eigen_GPU_warmnet_*.

Test infrastructure only: no product path imports it.  The reference has no
learned component; the network is the one the library defines
(eig_warmnet.hpp): features x = upper triangle (column by column) of
(A - mu I) / s with mu = trace(A) / n and s the RMS of that triangle,
h1 = relu(W1^T x + b1), h2 = relu(W2^T h1 + b2), y = W3^T h2 + b3 (n^2 values,
an n x n column-major matrix), loss = mean over the batch and the n^2 outputs
of (y - sign-fixed target)^2, each target column negated when its entry of
largest magnitude is negative.  Parameters: one vector, W1 (Din x H), b1,
W2 (H x H), b2, W3 (H x n^2), b3, matrices column-major.
"""
from __future__ import annotations

import numpy as np


def dims(n: int, h: int):
    din, dout = n * (n + 1) // 2, n * n
    sizes = [din * h, h, h * h, h, h * dout, dout]
    offs = np.cumsum([0] + sizes)
    return din, dout, offs


def unpack(p: np.ndarray, n: int, h: int):
    din, dout, o = dims(n, h)
    W1 = p[o[0]:o[1]].reshape(h, din).T
    b1 = p[o[1]:o[2]]
    W2 = p[o[2]:o[3]].reshape(h, h).T
    b2 = p[o[3]:o[4]]
    W3 = p[o[4]:o[5]].reshape(dout, h).T
    b3 = p[o[5]:o[6]]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([W1.T.ravel(), b1, W2.T.ravel(), b2, W3.T.ravel(), b3])


def features(A: np.ndarray) -> np.ndarray:
    """A: (B, n, n) symmetric (row-major).  Returns (B, Din)."""
    B, n, _ = A.shape
    mu = np.trace(A, axis1=1, axis2=2) / n
    M = A - mu[:, None, None] * np.eye(n)
    iu = [(i, j) for j in range(n) for i in range(j + 1)]
    rows = np.array([i for i, _ in iu])
    cols = np.array([j for _, j in iu])
    x = M[:, rows, cols]
    rms = np.sqrt(np.mean(x * x, axis=1))
    rms = np.where(rms > 0, rms, 1.0)
    return x / rms[:, None]


def forward(p, A, n, h):
    W1, b1, W2, b2, W3, b3 = unpack(p, n, h)
    x = features(A)
    z1 = x @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3 + b3
    return x, h1, h2, y


def sign_fixed(V: np.ndarray) -> np.ndarray:
    """V: (B, n, n) with eigenvector j in column j.  Returns (B, n*n): column-major flattening."""
    B, n, _ = V.shape
    idx = np.argmax(np.abs(V), axis=1)
    big = np.take_along_axis(V, idx[:, None, :], axis=1)[:, 0, :]
    T = V * np.where(big < 0, -1.0, 1.0)[:, None, :]
    return np.swapaxes(T, 1, 2).reshape(B, n * n)


def loss_and_grad(p, A, V, n, h):
    """Mean squared error and its gradient (the parameters' layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n, h)
    x, h1, h2, y = forward(p, A, n, h)
    T = sign_fixed(V)
    B = A.shape[0]
    d = y - T
    loss = float(np.sum(d * d) / (B * n * n))
    dy = 2.0 * d / (B * n * n)
    gW3 = h2.T @ dy
    gb3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (h2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1 > 0)
    gW1 = x.T @ dh1
    gb1 = dh1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)
