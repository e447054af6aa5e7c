"""CPU restatement (float64 numpy) of the library's eigenpair refinement step.

Test infrastructure only: the GPU tests compare eigen_GPU_batch_refine_DP
with this on the same inputs; no product path imports it.

The reference has no refinement (it lists "Half precision version + Iterative
refinment" as future work, README:169); this restates the published method the
library implements: Ogita & Aishima, "Iterative refinement for symmetric
eigenvalue decomposition", Japan J. Indust. Appl. Math. 35 (2018), algorithm 1,
with clusters resolved by Rayleigh-Ritz as in the clustered-eigenvalue variant
of the same authors (2019).  One step, for A symmetric and X approximate
eigenvectors:

  S = X^T A X, G = X^T X (symmetric: the upper triangle, mirrored), R = I - G
  l_i = s_ii / g_ii,  delta = 2 (||S - diag(l)||_F + max|l| ||R||_F)
  ascending order of l (ties by index); adjacent a, b join one cluster when
  l_b - l_a <= delta or max(|s_ab - l_b g_ab|, |s_ab - l_a g_ab|) > TAU (l_b - l_a)
  E_ij = (s_ij - l_j g_ij) / (l_j - l_i) between clusters, E_ii = (1 - g_ii)/2,
  inside a cluster of 2..KMAX members J (K the other indices): C = I + R_JJ / 2,
  F_ia = (s_ia - l_a g_ia) / (l_a - l_i) for i in K, the basis X_J C + X_K F's
  projected matrices S' = C^T S_JJ C + S_JK F + F^T S_KJ + F^T diag(l_K) F,
  G' = C^T G_JJ C + G_JK F + F^T G_KJ + F^T F, H = G' - I,
  P = I - H/2 + 3/8 H^2, W the (ascending) eigenvectors of P S' P:
  E_JJ = C P W - I, E_KJ = F P W; inside a larger cluster E_ij = -g_ij / 2
  X' = X + X E;  refined eigenvalues l (Ritz values inside resolved clusters)
"""
from __future__ import annotations

import numpy as np

KMAX = 8
TAU = 1e-2


def refine_step(A: np.ndarray, X: np.ndarray):
    """One step.  Returns (X', lnew, emax, big) where emax is the largest
    first-order entry (outside the resolved clusters' own blocks) and big
    tells whether a cluster exceeded KMAX."""
    n = A.shape[0]
    S = X.T @ A @ X
    G = X.T @ X
    S = np.triu(S) + np.triu(S, 1).T
    G = np.triu(G) + np.triu(G, 1).T
    lam = np.diag(S) / np.diag(G)
    R = np.eye(n) - G
    delta = 2.0 * (np.linalg.norm(S - np.diag(lam)) + np.abs(lam).max() * np.linalg.norm(R))
    order = np.lexsort((np.arange(n), lam))
    clusters = []
    p = 0
    while p < n:
        e = p + 1
        while e < n:
            ia, ib = order[e - 1], order[e]
            gap = lam[ib] - lam[ia]
            s, g = S[ia, ib], G[ia, ib]
            if gap <= delta or max(abs(s - lam[ib] * g), abs(s - lam[ia] * g)) > TAU * gap:
                e += 1
            else:
                break
        clusters.append(order[p:e])
        p = e
    cid = np.empty(n, dtype=int)
    for c, members in enumerate(clusters):
        cid[members] = c
    denom = lam[None, :] - lam[:, None]
    same = cid[:, None] == cid[None, :]
    F = np.where(same, 0.0, (S - lam[None, :] * G) / np.where(same, 1.0, denom))
    E = F.copy()
    np.fill_diagonal(E, 0.5 * (1.0 - np.diag(G)))
    lnew = lam.copy()
    big = False
    first_order = ~same
    for members in clusters:
        k = len(members)
        if k < 2:
            continue
        J = np.asarray(members)
        if k > KMAX:
            big = True
            for i in J:
                for j in J:
                    if i != j:
                        E[i, j] = -0.5 * G[i, j]
                        first_order[i, j] = True
            continue
        C = np.eye(k) + 0.5 * (np.eye(k) - G[np.ix_(J, J)])
        K = np.setdiff1d(np.arange(n), J)
        FK = (S[np.ix_(K, J)] - lam[J][None, :] * G[np.ix_(K, J)]) / (lam[J][None, :] - lam[K][:, None])
        Sp = C.T @ S[np.ix_(J, J)] @ C + S[np.ix_(J, K)] @ FK + FK.T @ S[np.ix_(K, J)] + (FK.T * lam[K]) @ FK
        Gp = C.T @ G[np.ix_(J, J)] @ C + G[np.ix_(J, K)] @ FK + FK.T @ G[np.ix_(K, J)] + FK.T @ FK
        H = Gp - np.eye(k)
        P = np.eye(k) - 0.5 * H + 0.375 * H @ H
        mu, W = np.linalg.eigh(P @ Sp @ P)
        PW = P @ W
        E[np.ix_(K, J)] = FK @ PW
        E[np.ix_(J, J)] = C @ PW - np.eye(k)
        lnew[J] = mu
    emax = float(np.abs(np.where(first_order, E, 0.0)).max()) if n > 1 else 0.0
    return X + X @ E, lnew, emax, big


def refine(A: np.ndarray, X: np.ndarray, iters: int):
    """iters steps; returns (X sorted by eigenvalue, eigenvalues ascending, emax of the last step)."""
    lnew, emax = None, 0.0
    for _ in range(iters):
        X, lnew, emax, _ = refine_step(A, X)
    order = np.lexsort((np.arange(len(lnew)), lnew))
    return X[:, order], lnew[order], emax
