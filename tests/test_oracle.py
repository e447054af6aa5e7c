"""Pins the CPU oracle (oracle/eispack_impl.h) before it is trusted as the checker.

Every check here is independent of the oracle's own formulas: LAPACK through
numpy (dsyevd/ssyevd), closed forms (Frank, Helmert-LETKF, 2x2, diagonal),
invariants (A = Q T Q^T from the stored Householder vectors, Z^T Z = I,
T Z = Z D), a hand-derived tie case of the reference's selection sort, and the
reference's own published checker thresholds (logs/log-MI100:202,
logs/log-RX480:226).  Planted mistakes in copies of the oracle's output must
trip the checker.
"""
import numpy as np
import pytest

import oracle
import workloads as wl


def eps_of(dt):
    return float(np.finfo(dt).eps)


def ql_tol(dt):
    """The reference deflates |e(m)| <= tol*(|d(m)|+|d(m+1)|) with tol = 512 eps
    (FP64) / 16 eps (FP32) (tql2_tiled.hpp:8), a first-order residual by design."""
    return (512.0 if np.dtype(dt) == np.float64 else 16.0) * eps_of(dt)


def solve(a, solver=oracle.IMTQL2):
    z, w, info = oracle.eig_batch(np.nan_to_num(a), solver)
    assert info.max() == 0
    return z, w


@pytest.mark.parametrize("solver", [oracle.IMTQL2, oracle.TQL2, oracle.JACOBI, oracle.JACOBI_CLASSICAL])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 16, 31, 32, 33, 47, 64, 65, 100, 128])
def test_eigenvalues_vs_lapack(n, dtype, solver):
    nm = n + (n & 1)
    a = wl.sym_rand(6, n, nm, dtype=dtype, seed=n)
    z, w = solve(a, solver)
    full = wl.unpack(np.nan_to_num(a)).astype(np.float64)
    ref = np.linalg.eigvalsh(full)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(w - ref) / scale
    assert err.max() <= 2 * ql_tol(dtype) + 8 * n * eps_of(dtype)
    # eigenpairs: per-pair residual and orthogonality in double
    Z = np.swapaxes(z[:, :, :n], 1, 2).astype(np.float64)        # Z[k, r, j]
    res = np.einsum("krc,kcj->krj", full, Z) - Z * w[:, None, :].astype(np.float64)
    assert np.linalg.norm(res, axis=1).max() / scale.max() <= 2 * ql_tol(dtype) + 16 * n * eps_of(dtype)
    orth = np.einsum("krj,kri->kji", Z, Z) - np.eye(n)
    assert np.abs(orth).max() <= 16 * n * eps_of(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 32, 48, 64, 129])
def test_frank_closed_form(n, dtype):
    a = wl.frank(2, n, n + (n & 1), dtype=dtype)
    _, w = solve(a)
    lam = wl.frank_eigenvalues(n)
    assert np.max(np.abs(w - lam) / lam[-1]) <= 2 * ql_tol(dtype) + 8 * n * eps_of(dtype)


def test_frank_closed_form_is_what_the_reference_checks():
    # eigen_GPU_check.cu:17-32 lists 1/(2(1-cos((2(n+1-i)-1) pi/(2n+1)))) for
    # i = 1..n, which is the same set as frank_eigenvalues; brute force n <= 6.
    for n in range(1, 7):
        i = np.arange(1, n + 1)
        ref = 1.0 / (2 * (1 - np.cos((2 * (n + 1 - i) - 1) * np.pi / (2 * n + 1))))
        A = np.minimum.outer(np.arange(1, n + 1), np.arange(1, n + 1)).astype(float)
        assert np.allclose(np.sort(ref), np.linalg.eigvalsh(A), rtol=1e-13)
        assert np.allclose(np.sort(ref), wl.frank_eigenvalues(n), rtol=1e-15)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_letkf_helmert_closed_form(dtype):
    n = 40
    a, lam = wl.letkf_helmert(4, n, n, dtype=dtype, seed=2)
    z, w = solve(a)
    assert np.max(np.abs(w - lam) / lam[:, -1:]) <= 2 * ql_tol(dtype) + 8 * n * eps_of(dtype)
    H = wl.helmert(n)
    assert np.allclose(H @ H.T, np.eye(n), atol=1e-14)


def test_two_by_two_closed_form():
    rng = np.random.default_rng(9)
    for _ in range(50):
        p, q, r = rng.normal(size=3)
        A = np.array([[p, q], [q, r]])
        a = np.zeros((1, 2, 2))
        a[0] = A.T
        z, w = solve(a)
        mid, rad = (p + r) / 2, np.hypot((p - r) / 2, q)
        assert np.allclose(w[0], [mid - rad, mid + rad], rtol=1e-14, atol=1e-14)
        Z = z[0].T
        assert np.allclose(A @ Z, Z * w[0], atol=1e-13)


def test_one_by_one():
    a = np.array([[[3.5]]])
    z, w = solve(a)
    assert w[0, 0] == 3.5 and z[0, 0, 0] == 1.0


def test_sort_tie_rule_hand_derived():
    # diag(2, 1, 2, 1): the reduction and QL leave it untouched, so only the
    # reference's selection sort (tql2_tiled.hpp:149-176) orders the vectors:
    # l=0 swaps slots 0,1 -> d=[1,2,2,1] pos=[1,0,2,3]; l=1 swaps slots 1,3 ->
    # d=[1,1,2,2] pos=[1,3,2,0]; l=2 finds no strictly smaller -> no swap.
    D = np.diag([2.0, 1.0, 2.0, 1.0])
    a = D.T[None].copy()
    z, w = solve(a)
    assert np.array_equal(w[0], [1.0, 1.0, 2.0, 2.0])
    cols = [int(np.argmax(np.abs(z[0, j]))) for j in range(4)]
    assert cols == [1, 3, 2, 0]
    assert np.allclose(np.abs(z[0]), np.eye(4)[cols])


def test_tridiagonal_storage_convention():
    # a(j,j) = e(j-1), u in a(1:l, i), beta_i = 1/(a(i,i) a(l,i)): rebuilding
    # Q = H_n ... H_2 from the storage must give A = Q T Q^T (tred1_tiled.hpp:207-212).
    rng = np.random.default_rng(4)
    for n in (2, 3, 8, 17):
        B = rng.normal(size=(n, n))
        A = B + B.T
        st, d, e = oracle.tred1(A.T.copy())          # column-major storage = A.T in C order
        S = st.T                                       # S[r, c] = a(r+1, c+1)
        T = np.diag(d) + np.diag(e[:-1], 1) + np.diag(e[:-1], -1)
        assert e[-1] == 0.0
        Q = np.eye(n)
        for i in range(n, 1, -1):                      # creation order i = n..2
            l = i - 1
            u = np.zeros(n)
            u[:l] = S[:l, i - 1]
            prod = S[i - 1, i - 1] * S[l - 1, i - 1]
            beta = 1.0 / (prod if prod != 0 else 1.0)
            assert np.isclose(S[i - 1, i - 1], e[l - 1])
            Q = Q @ (np.eye(n) + beta * np.outer(u, u))
        assert np.allclose(Q.T @ Q, np.eye(n), atol=1e-13)
        assert np.allclose(Q @ T @ Q.T, A, atol=1e-12 * np.abs(A).max())


@pytest.mark.parametrize("variant", [oracle.IMTQL2, oracle.TQL2])
def test_ql_on_tridiagonal(variant):
    rng = np.random.default_rng(11 + variant)
    for n in (1, 2, 5, 30, 64):
        d0 = rng.normal(size=n)
        e0 = np.zeros(n)
        e0[:-1] = rng.normal(size=n - 1)
        T = np.diag(d0) + np.diag(e0[:-1], 1) + np.diag(e0[:-1], -1)
        d, z, rc = oracle.ql(d0, e0, variant)
        assert rc == 0
        Z = z.T                                     # Z[r, j]
        assert np.allclose(np.sort(d), np.linalg.eigvalsh(T), atol=1e-13)
        assert np.allclose(T @ Z, Z * d, atol=1e-12)
        assert np.allclose(Z.T @ Z, np.eye(n), atol=1e-13)


def test_checker_thresholds_match_reference_logs():
    # TOL printed by the reference checker: 3.305345e-05 (FP32, n=64,
    # logs/log-MI100:202) and 2.031706e-12 (FP64, n=64, logs/log-RX480:226).
    for dt, tol in ((np.float32, 3.305345e-05), (np.float64, 2.031706e-12)):
        a = wl.frank(1, 64, 64, dtype=dt)
        z, w = solve(a)
        out = oracle.check(np.nan_to_num(a[0]), w[0], z[0])
        assert abs(out[2] - tol) / tol < 1e-6
        assert out[3] == 0


def test_checker_catches_planted_mistakes():
    n = 24
    a = wl.sym_rand(1, n, n, dtype=np.float64, seed=1)
    z, w = solve(a)
    good = oracle.check(a[0], w[0], z[0])
    assert good[3] == 0
    zb = z[0].copy()
    zb[5] *= -1.0                                   # sign flip: still an eigenpair
    assert oracle.check(a[0], w[0], zb)[3] == 0
    zb = z[0].copy()
    zb[[3, 4]] = zb[[4, 3]]                         # swapped vectors: residual fails
    assert int(oracle.check(a[0], w[0], zb)[3]) & 1
    zb = z[0].copy()
    zb[7] *= 1.0 + 1e-6                             # broken normalisation: orthogonality fails
    assert int(oracle.check(a[0], w[0], zb)[3]) & 2
    wb = w[0].copy()
    wb[0] += 1e-9                                    # wrong eigenvalue: residual fails
    assert int(oracle.check(a[0], wb, z[0])[3]) & 1


def test_nonfinite_input_reports_error():
    a = wl.sym_rand(3, 8, 8, dtype=np.float64, seed=3)
    a[1, 2, 3] = np.nan
    a[1, 3, 2] = np.nan
    _, _, info = oracle.eig_batch(a, oracle.IMTQL2)
    assert info[0] == 0 and info[2] == 0 and info[1] > 0


def test_invalid_arguments():
    with pytest.raises(ValueError):
        oracle.eig_batch(np.zeros((0, 4, 4)))


def test_workload_generators():
    a, Y = wl.letkf_y(3, k=16, p=40, rho=1.1, sigma=0.1, dtype=np.float64, seed=0)
    A = wl.unpack(a)
    assert np.allclose(A, np.swapaxes(A, 1, 2))
    lam = np.linalg.eigvalsh(A)
    assert lam.min() >= 15 / 1.1 - 1e-12
    lam0 = wl.clustered_uniform(2, 40, seed=1)
    q = wl.qdq(lam0, dtype=np.float64, seed=2)
    assert np.allclose(np.linalg.eigvalsh(wl.unpack(q)), np.sort(lam0, axis=1), atol=1e-12)
    ls = wl.logspaced_shuffled(2, 10)
    assert np.isclose(ls.max(), 1.0) and np.isclose(ls.min(), 1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_classical_jacobi_special_cases(dtype):
    """The classical (jac.cpp) Jacobi: a diagonal matrix needs no rotation
    (the guard for all-zero column maxima), the identity keeps Z = I, the
    2 x 2 closed form, and the Frank closed form."""
    n = 9
    d = np.array([3.0, -1.0, 2.0, 2.0, 0.5, 7.0, -4.0, 0.0, 1.0])
    a = np.broadcast_to(np.diag(d), (1, n, n)).astype(dtype).copy()
    z, w = solve(a, oracle.JACOBI_CLASSICAL)
    assert np.array_equal(w[0], np.sort(d).astype(dtype))
    assert np.array_equal(np.abs(z[0]), np.abs(z[0]).round())           # a permutation
    a2 = np.array([[[2.0, 1.0], [1.0, 2.0]]], dtype=dtype)              # eigenvalues 1, 3
    z, w = solve(a2, oracle.JACOBI_CLASSICAL)
    assert np.allclose(w[0], [1.0, 3.0], rtol=4 * eps_of(dtype))
    assert np.allclose(np.abs(z[0]), np.sqrt(0.5), rtol=8 * eps_of(dtype))
    af = wl.frank(1, 24, 24, dtype=dtype)
    _, w = solve(af, oracle.JACOBI_CLASSICAL)
    lam = wl.frank_eigenvalues(24)
    assert np.max(np.abs(w - lam) / lam[-1]) <= 64 * 24 * eps_of(dtype)


def test_refinement_restatement_converges_to_lapack():
    """oracle/refine_oa.py (the restatement the GPU refinement is checked
    against) is pinned by what the mathematics fixes: from FP32 eigenvectors,
    three steps reach FP64 eigenpairs of LAPACK's accuracy (eigenvalues equal to
    numpy's eigvalsh within 1e-13 max|l|, residual and orthogonality ~1e-14),
    and clusters below FP32 resolution (pairs split by 1e-8) come out resolved."""
    import refine_oa
    for n in (2, 7, 48, 120):
        lam = wl.clustered_uniform(3, n, frac=0.1, gap=1e-8, seed=n)
        full = wl.unpack(wl.qdq(lam, dtype=np.float64, seed=n + 1))
        for k in range(3):
            A = full[k]
            _, v = np.linalg.eigh(A.astype(np.float32))
            X = v.astype(np.float64)
            ems = []
            for _ in range(3):
                X, l, em, big = refine_oa.refine_step(A, X)
                ems.append(em)
                assert not big
            order = np.argsort(l)
            z, w = X[:, order], l[order]
            ref = np.linalg.eigvalsh(A)
            assert np.max(np.abs(w - ref)) <= 1e-13 * np.abs(ref).max()
            assert np.linalg.norm(A @ z - z * w) <= 1e-13 * np.abs(w).max()
            assert np.linalg.norm(z.T @ z - np.eye(n)) <= 1e-13
            assert ems[2] <= 1e-6          # the GPU's acceptance bound on the last step


def test_refinement_restatement_catches_a_planted_mistake(monkeypatch):
    """A sign error in the first-order formula (E_ij with l_i - l_j) no longer
    converges: the check above would fail."""
    import refine_oa
    n = 30
    full = wl.unpack(wl.qdq(wl.clustered_uniform(1, n, seed=4), dtype=np.float64, seed=5))
    A = full[0]
    _, v = np.linalg.eigh(A.astype(np.float32))
    X = v.astype(np.float64)
    orig = refine_oa.refine_step

    def bad_step(A, X):
        Xn, l, em, big = orig(A, X)
        # the same step with E's off-diagonal first-order part negated
        E = np.linalg.lstsq(X, Xn, rcond=None)[0] - np.eye(len(l))
        D = np.diag(np.diag(E))
        return X + X @ (D - (E - D)), l, em, big
    for _ in range(3):
        X, l, _, _ = bad_step(A, X)
    assert np.linalg.norm(A @ X - X * l) > 1e-10


def test_warmnet_reference_gradient_is_the_derivative():
    """oracle/warmnet_ref.py (the checker of the GPU network kernels) is pinned
    by central differences of its own loss (every gradient entry agrees) and by
    the loss being the mean squared distance to the sign-fixed eigenvectors."""
    import warmnet_ref
    rng = np.random.default_rng(0)
    n, h, B = 3, 5, 4
    din, dout, offs = warmnet_ref.dims(n, h)
    p = rng.normal(size=offs[-1]) * 0.5
    G = rng.normal(size=(B, n, n))
    A = G + np.swapaxes(G, 1, 2)
    V = np.linalg.eigh(A)[1]
    loss, g = warmnet_ref.loss_and_grad(p, A, V, n, h)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = eps
        fd[i] = (warmnet_ref.loss_and_grad(p + e, A, V, n, h)[0] - warmnet_ref.loss_and_grad(p - e, A, V, n, h)[0]) / (2 * eps)
    assert np.max(np.abs(fd - g)) <= 1e-6 * max(1.0, np.abs(g).max())
    # the loss is the mean squared distance to the sign-fixed eigenvectors
    _, _, _, y = warmnet_ref.forward(p, A, n, h)
    T = warmnet_ref.sign_fixed(V)
    assert abs(loss - np.mean((y - T) ** 2)) <= 1e-14
    assert np.all(np.abs(T.reshape(B, n, n)).argmax(axis=2) >= 0)
    Tm = np.swapaxes(T.reshape(B, n, n), 1, 2)
    big = np.take_along_axis(Tm, np.abs(Tm).argmax(axis=1)[:, None, :], axis=1)
    assert np.all(big > 0)
