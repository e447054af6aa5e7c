"""Synthetic input families (numpy, seeded) in the reference's batch layout.

Layout: array of shape (L, n, nm), ``a[k, c, r]`` = element (r, c) of matrix k
(column-major with leading dimension nm, matrix stride nm*n, as
eigen_GPU_batch.cu:74-90).  Padding rows r >= n are filled with NaN so a kernel
that reads them is caught by the tests.

Families:
  frank          a(i,j) = min(i,j)                       main.cpp:182-187
  sym_rand       A = (B + B^T)/2, B_ij ~ U[-1,1]          main.cpp:175-181 (as specified
                                                          for the n=32 FP64 parity case)
  letkf_helmert  A = H^T diag(w) H, Helmert basis H,       main.cpp:68-167
                 w_i = (n-1)(1+delta_i) + 2 random spikes
  letkf_y        A = ((k-1)/rho) I + Y^T R^-1 Y, R = I,   LETKF ensemble-space matrix
                 Y (p x k) ~ N(0, sigma^2)
  qdq            A = Q diag(lam) Q^T, Q from the QR of a Gaussian matrix
The closed forms used by the tests: Frank eigenvalues 1/(2-2cos((2j-1)pi/(2n+1)))
(eigen_GPU_check.cu:17-32), Helmert-LETKF eigenpairs (w, Helmert rows), and the
prescribed lam of qdq.
"""
from __future__ import annotations

import numpy as np


def _pack(A: np.ndarray, nm: int, dtype) -> np.ndarray:
    """(L, n, n) symmetric (row-major A[k, r, c]) -> (L, n, nm) column-major batch."""
    L, n, _ = A.shape
    out = np.full((L, n, nm), np.nan, dtype=dtype)
    out[:, :, :n] = np.swapaxes(A, 1, 2)
    return out


def unpack(a: np.ndarray) -> np.ndarray:
    """(L, n, nm) batch -> (L, n, n) matrices M[k, r, c]."""
    n = a.shape[1]
    return np.swapaxes(a[:, :, :n], 1, 2)


def frank(L: int, n: int, nm: int | None = None, dtype=np.float64) -> np.ndarray:
    nm = n if nm is None else nm
    i = np.arange(1, n + 1)
    A = np.minimum.outer(i, i).astype(np.float64)
    return _pack(np.broadcast_to(A, (L, n, n)), nm, dtype)


def frank_eigenvalues(n: int) -> np.ndarray:
    j = np.arange(1, n + 1)
    lam = 1.0 / (2.0 - 2.0 * np.cos((2 * j - 1) * np.pi / (2 * n + 1)))
    return np.sort(lam)


def sym_rand(L: int, n: int, nm: int | None = None, dtype=np.float64, seed: int = 0) -> np.ndarray:
    nm = n if nm is None else nm
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1.0, 1.0, size=(L, n, n))
    A = 0.5 * (B + np.swapaxes(B, 1, 2))
    return _pack(A, nm, dtype)


def helmert(n: int) -> np.ndarray:
    """Rows h_i of the Helmert orthogonal matrix (main.cpp:112-127)."""
    H = np.zeros((n, n))
    H[0, :] = 1.0 / np.sqrt(n)
    for i in range(2, n + 1):
        s = np.sqrt((i - 1) * i)
        H[i - 1, : i - 1] = 1.0 / s
        H[i - 1, i - 1] = -(i - 1) / s
    return H


def letkf_helmert(L: int, n: int, nm: int | None = None, dtype=np.float64, seed: int = 0):
    """Reference LETKF-like family; returns (batch, eigenvalues (L, n) ascending)."""
    nm = n if nm is None else nm
    rng = np.random.default_rng(seed)
    eps = np.finfo(dtype).eps
    H = helmert(n)
    A = np.empty((L, n, n))
    lam = np.empty((L, n))
    for k in range(L):
        w = (n - 1) * (1.0 + 15.0 * (2.0 * rng.uniform(size=n) - 1.0) * eps)
        for _ in range(2):
            w[rng.integers(n)] += rng.uniform() * n
        A[k] = (H.T * w) @ H
        lam[k] = np.sort(w)
    return _pack(A, nm, dtype), lam


def letkf_y(L: int, k: int = 64, p: int = 400, rho: float = 1.1, sigma: float = 0.1,
            nm: int | None = None, dtype=np.float32, seed: int = 0, Y: np.ndarray | None = None):
    """A = ((k-1)/rho) I + Y^T Y (R = I); returns (batch, Y)."""
    nm = k if nm is None else nm
    rng = np.random.default_rng(seed)
    if Y is None:
        Y = rng.normal(0.0, sigma, size=(L, p, k))
    A = np.einsum("lpi,lpj->lij", Y, Y)
    A[:, np.arange(k), np.arange(k)] += (k - 1) / rho
    return _pack(A, nm, dtype), Y


def qdq(lam: np.ndarray, nm: int | None = None, dtype=np.float64, seed: int = 0) -> np.ndarray:
    """A_k = Q_k diag(lam_k) Q_k^T with Q_k from the QR of a Gaussian matrix."""
    L, n = lam.shape
    nm = n if nm is None else nm
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(L, n, n))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
    A = (Q * lam[:, None, :]) @ np.swapaxes(Q, 1, 2)
    A = 0.5 * (A + np.swapaxes(A, 1, 2))
    return _pack(A, nm, dtype)


def logspaced_shuffled(L: int, n: int, hi: float = 1.0, lo: float = 1e-6, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    base = np.logspace(np.log10(hi), np.log10(lo), n)
    return np.stack([rng.permutation(base) for _ in range(L)])


def clustered_uniform(L: int, n: int, frac: float = 0.05, gap: float = 1e-8, seed: int = 0) -> np.ndarray:
    """lam ~ U[-1,1] with a fraction of them replaced by pairs split by `gap`."""
    rng = np.random.default_rng(seed)
    lam = rng.uniform(-1.0, 1.0, size=(L, n))
    npairs = max(1, int(round(frac * n / 2)))
    for k in range(L):
        idx = rng.choice(n, size=2 * npairs, replace=False)
        lam[k, idx[1::2]] = lam[k, idx[0::2]] + gap
    return lam
