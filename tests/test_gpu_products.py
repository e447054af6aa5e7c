"""GPU tests of the MFMA products around the solver:

* symdmt  C = Z diag(d) Z^T      (eigen_GPU_batch_symdmt_*; the reference's symdmt.cu:5-87)
* matpow  C = s Z diag(w)^p Z^T  (eigen_GPU_batch_matpow_*; README:171)
* gram    A = alpha I + Y^T diag(r) Y  (eigen_GPU_batch_gram_*; the LETKF ingest)

Floating-point kernels, so each is compared with an FP64 numpy product of the
same inputs, element by element, against the dot-product error bound
|C - C_ref| <= c (m + 2) eps (|X| |D| |X|^T) (m the contraction length, c = 3,
tolerance written in check_product).  The outputs must be exactly symmetric
and must leave the padding rows n..nm-1 untouched.
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg

pytestmark = pytest.mark.gpu

DEV = "cuda"
EPS = {np.float32: np.finfo(np.float32).eps, np.float64: np.finfo(np.float64).eps}


def check_product(c_gpu, x, dvec, m, dtype, diag=0.0):
    """c_gpu (n, n) vs x diag(dvec) x^T + diag I in FP64, with the bound above."""
    x64 = x.astype(np.float64)
    d64 = dvec.astype(np.float64)
    ref = (x64 * d64) @ x64.T + diag * np.eye(x.shape[0])
    bound = 3.0 * (m + 2) * EPS[dtype] * ((np.abs(x64) * np.abs(d64)) @ np.abs(x64).T + abs(diag) * np.eye(x.shape[0]))
    err = np.abs(c_gpu.astype(np.float64) - ref)
    assert np.all(err <= bound + np.finfo(dtype).tiny), f"max excess {np.max(err - bound):.3e}"
    np.testing.assert_array_equal(c_gpu, c_gpu.T)


SIZES = [1, 3, 16, 31, 32, 33, 47, 64, 65, 96, 100, 128, 200, 256, 300, 512]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", SIZES)
def test_symdmt_vs_fp64(n, dtype):
    print(f"[products] symdmt n={n} dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(100 + n)
    L = 6 if n <= 256 else 3
    nm = n + (n & 1)
    z = np.full((L, n, nm), np.nan, dtype=dtype)
    z[:, :, :n] = rng.standard_normal((L, n, n))
    d = rng.uniform(-2.0, 2.0, size=(L, n)).astype(dtype)
    zt, dt = torch.from_numpy(z).to(DEV), torch.from_numpy(d).to(DEV)
    c = torch.full_like(zt, np.nan)
    eg.symdmt(zt, dt, c)
    torch.cuda.synchronize()
    c = c.cpu().numpy()
    if nm > n:
        assert np.isnan(c[:, :, n:]).all()
    for k in range(L):
        # Z(r, c) = z[k, c, r]
        check_product(c[k, :, :n], z[k, :, :n].T, d[k], n, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [5, 48, 64, 100, 256])
def test_matpow_from_solver(n, dtype):
    """Solve SPD matrices, then: p = 1 gives A back, p = -1 its inverse, p = 0.5
    a square root, p = -0.5 with scale s the matrix s A^-1/2."""
    print(f"[products] matpow n={n} dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(7 + n)
    L = 4
    B = rng.standard_normal((L, n, n))
    A = B @ np.swapaxes(B, 1, 2) / n + np.eye(n)            # SPD, condition O(10)
    a = torch.from_numpy(np.swapaxes(A, 1, 2).astype(dtype).copy()).to(DEV)
    z, w, info = eg.eigen_batch(a.clone())
    assert int(info.abs().max()) == 0
    tol = 200 * n * EPS[dtype]
    A64 = A.astype(dtype).astype(np.float64)
    for p, s in ((1.0, 1.0), (-1.0, 1.0), (0.5, 1.0), (-0.5, 3.0)):
        c = eg.matpow(z, w, p, s)
        torch.cuda.synchronize()
        C = np.swapaxes(c.cpu().numpy(), 1, 2).astype(np.float64)
        for k in range(L):
            np.testing.assert_array_equal(C[k], C[k].T)
            nrm = np.linalg.norm(A64[k])
            if p == 1.0:
                err = np.linalg.norm(C[k] - A64[k]) / nrm
            elif p == -1.0:
                err = np.linalg.norm(C[k] @ A64[k] - np.eye(n)) / np.sqrt(n)
            elif p == 0.5:
                err = np.linalg.norm(C[k] @ C[k] - A64[k]) / nrm
            else:
                err = np.linalg.norm(C[k] @ C[k] @ A64[k] / (s * s) - np.eye(n)) / np.sqrt(n)
            assert err <= tol, f"p={p} matrix {k}: {err:.3e} > {tol:.3e}"


GRAM_CASES = [  # (n, p, ldy extra rows, rinv mode)
    (1, 1, 0, "none"), (17, 7, 1, "shared"), (48, 300, 0, "none"), (64, 400, 0, "per"),
    (64, 401, 3, "per"), (64, 16, 0, "shared"), (100, 33, 2, "none"), (256, 64, 0, "per"),
    (300, 20, 4, "shared"), (512, 8, 0, "none"),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,p,extra,mode", GRAM_CASES)
def test_gram_vs_fp64(n, p, extra, mode, dtype):
    print(f"[products] gram n={n} p={p} ldy=p+{extra} rinv={mode} dtype={np.dtype(dtype).name}", flush=True)
    rng = np.random.default_rng(n * 1000 + p)
    L = 5 if n <= 256 else 2
    ldy = p + extra
    nm = n + 1
    y = np.full((L, n, ldy), np.nan, dtype=dtype)
    y[:, :, :p] = rng.normal(0.0, 0.1, size=(L, n, p))
    alpha = 63.0 / 1.1
    if mode == "none":
        r, rt, rstride = np.ones((L, p)), None, 0
    elif mode == "shared":
        r1 = rng.uniform(0.5, 2.0, size=p).astype(dtype)
        r, rt, rstride = np.broadcast_to(r1, (L, p)), torch.from_numpy(r1).to(DEV), 0
    else:
        rstride = p + (p % 4 and 4 - p % 4)
        rr = np.zeros((L, rstride), dtype=dtype)
        rr[:, :p] = rng.uniform(0.5, 2.0, size=(L, p))
        r, rt = rr[:, :p], torch.from_numpy(rr.reshape(-1).copy()).to(DEV)
    a = torch.full((L, n, nm), np.nan, dtype=torch.float32 if dtype == np.float32 else torch.float64, device=DEV)
    eg.gram(torch.from_numpy(y).to(DEV), p, alpha, rinv=rt, rstride=rstride, nm=nm, a=a)
    torch.cuda.synchronize()
    a = a.cpu().numpy()
    assert np.isnan(a[:, :, n:]).all()
    for k in range(L):
        Yk = y[k, :, :p].T                                    # p x n
        check_product(a[k, :, :n], Yk.T, np.asarray(r[k], dtype=dtype), p, dtype, diag=alpha)


def test_letkf_chain_on_gpu():
    """The whole LETKF analysis step on the GPU: A = (k-1)/rho I + Y^T Y, its
    eigenpairs, then T = sqrt(k-1) A^-1/2, against FP64 numpy."""
    rng = np.random.default_rng(3)
    L, k, p, rho = 32, 64, 400, 1.1
    y = rng.normal(0.0, 0.1, size=(L, k, p)).astype(np.float32)
    a = eg.gram(torch.from_numpy(y).to(DEV), p, (k - 1) / rho)
    z, w, info = eg.eigen_batch(a.clone())
    assert int(info.abs().max()) == 0
    t = eg.matpow(z, w, -0.5, float(np.sqrt(k - 1)))
    torch.cuda.synchronize()
    T = np.swapaxes(t.cpu().numpy(), 1, 2).astype(np.float64)
    for i in range(L):
        Y = y[i].T.astype(np.float64)
        A = (k - 1) / rho * np.eye(k) + Y.T @ Y
        lam, V = np.linalg.eigh(A)
        Tref = np.sqrt(k - 1) * (V / np.sqrt(lam)) @ V.T
        err = np.linalg.norm(T[i] - Tref) / np.linalg.norm(Tref)
        assert err <= 1e-5, f"matrix {i}: {err:.3e}"


def test_products_invalid_args():
    z = torch.zeros((2, 4, 4), device=DEV)
    d = torch.zeros((2, 4), device=DEV)
    with pytest.raises(RuntimeError):
        eg.symdmt(z, d, z)                                   # output overlaps input
    big = torch.zeros((1, 513, 513), device=DEV)
    with pytest.raises(RuntimeError):
        eg.symdmt(big, torch.zeros((1, 513), device=DEV))
    with pytest.raises(RuntimeError):
        eg.gram(torch.zeros((2, 4, 3), device=DEV), 5)         # p > ldy
    # L = 0 is a no-op
    rc = eg.lib().eigen_GPU_batch_symdmt_FP(0, 4, 4, z.data_ptr(), d.data_ptr(), 0, 0)
    assert rc != 0                                           # null output pointer
    c = torch.empty_like(z)
    rc = eg.lib().eigen_GPU_batch_symdmt_FP(0, 4, 4, z.data_ptr(), d.data_ptr(), c.data_ptr(), 0)
    assert rc == 0
