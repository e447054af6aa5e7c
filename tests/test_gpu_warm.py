"""GPU: the batched Householder QR (eigen_GPU_batch_qr_*) and the warm-started
solve (eigen_GPU_batch_warm_*: Q = qr(q0), B = Q^T A Q, capped Jacobi, V = Q Z_B,
plain fallback).

Bars: Q orthonormal (||Q^T Q - I||_F <= 8 n eps) and Q^T M upper triangular with
a non-negative diagonal (below-diagonal part <= 8 n eps ||M||_F), equal to
numpy's sign-fixed QR within 1e3 n eps; the warm-started eigenpairs pass the
reference checker (eigen_GPU_check.cu:34-81) and their eigenvalues equal the
plain solve's within its TOL; the Jacobi sweep count reflects the quality of q0.
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import workloads as wl

pytestmark = pytest.mark.gpu
DEV = "cuda"


def ref_tol(n, dtype):
    eps = 2.0 ** -52 if dtype == np.float64 else 2.0 ** -23
    return eps * (1023.0 if dtype == np.float64 else 31.0) * np.sqrt(n + 16.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 5, 33, 48, 64])
def test_qr(n, dtype):
    print(f"[qr] n={n} {np.dtype(dtype).name}", flush=True)
    L, nm = 20, n + 1
    rng = np.random.default_rng(n)
    M = rng.normal(size=(L, n, n))
    M[0, :, 0] = 0.0                                    # a zero column: no reflection there
    a = np.full((L, n, nm), np.nan)
    a[:, :, :n] = np.swapaxes(M, 1, 2)                  # column-major with ld nm
    at = torch.from_numpy(a.astype(dtype)).to(DEV)
    q = eg.qr(at, q=at.clone())                         # padding rows of q: NaN, left untouched
    torch.cuda.synchronize()
    Q = np.swapaxes(q.cpu().numpy()[:, :, :n], 1, 2).astype(np.float64)
    eps = np.finfo(dtype).eps
    Mr = M.astype(dtype).astype(np.float64)
    for k in range(L):
        assert np.linalg.norm(Q[k].T @ Q[k] - np.eye(n)) <= 8 * n * eps
        R = Q[k].T @ Mr[k]
        assert np.linalg.norm(np.tril(R, -1)) <= 8 * n * eps * max(np.linalg.norm(Mr[k]), 1.0)
        assert np.all(np.diag(R) >= -8 * n * eps * max(np.linalg.norm(Mr[k]), 1.0))
        if k > 0:
            Qn, Rn = np.linalg.qr(Mr[k])
            Qn = Qn * np.sign(np.diag(Rn))
            assert np.max(np.abs(Q[k] - Qn)) <= 1e3 * n * eps
    assert np.isnan(q.cpu().numpy()[:, :, n:]).all()


def solve_plain(a0):
    a = a0.clone()
    a, w, info = eg.eigen_batch(a)
    torch.cuda.synchronize()
    return a, w


def check_pairs(a0_np, z, w, dtype):
    L, n, nm = a0_np.shape
    full = wl.unpack(a0_np)
    for k in range(L):
        a_full = np.zeros((n, nm), dtype=dtype)
        a_full[:, :n] = full[k].T
        out = oracle.check(a_full, w[k], z[k])
        assert int(out[3]) == 0, (k, out[:3])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [8, 48, 64])
def test_warm_from_a_perturbed_basis(n, dtype):
    """q0 = the plain solve's eigenvectors plus noise of 1e-3: few sweeps, no
    fallback, eigenpairs to the plain solver's accuracy."""
    print(f"[warm] perturbed n={n} {np.dtype(dtype).name}", flush=True)
    L = 300
    a0_np, _ = wl.letkf_y(L, k=n, p=300, rho=1.05, dtype=dtype, seed=n)
    a0 = torch.from_numpy(a0_np).to(DEV)
    zp, wp = solve_plain(a0)
    g = torch.Generator(device="cpu").manual_seed(1)
    q0 = (zp.cpu() + 1e-3 * torch.randn(zp.shape, generator=g, dtype=zp.dtype)).to(DEV)
    a = a0.clone()
    a, w, info, sw = eg.eigen_batch_warm(a, q0, max_sweeps=8)
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0
    sw = sw.cpu().numpy()
    assert sw.max() < 8, sw.max()
    wn, z = w.cpu().numpy(), a.cpu().numpy()
    scale = np.abs(wn).max(axis=1, keepdims=True)
    assert np.max(np.abs(wn - wp.cpu().numpy()) / scale) <= ref_tol(n, dtype)
    check_pairs(a0_np, z, wn, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_warm_identity_start_and_fallback(dtype):
    """q0 = I is the plain Jacobi (more sweeps than from a good basis); with a
    one-sweep cap every matrix falls back to the plain solver, and the results
    still meet the bar; m < n writes the first m columns only."""
    print(f"[warm] identity / fallback {np.dtype(dtype).name}", flush=True)
    L, n = 200, 40
    a0_np = wl.sym_rand(L, n, n, dtype=dtype, seed=7)
    a0 = torch.from_numpy(a0_np).to(DEV)
    eye = torch.eye(n, dtype=a0.dtype, device=DEV).expand(L, n, n).contiguous()
    a, w, info, sw = eg.eigen_batch_warm(a0.clone(), eye, max_sweeps=30)
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0
    assert sw.cpu().numpy().min() >= 4
    check_pairs(a0_np, a.cpu().numpy(), w.cpu().numpy(), dtype)
    a1, w1, info1, sw1 = eg.eigen_batch_warm(a0.clone(), eye, max_sweeps=1)
    torch.cuda.synchronize()
    assert int(info1.abs().max()) == 0 and int(sw1.min()) == 1
    check_pairs(a0_np, a1.cpu().numpy(), w1.cpu().numpy(), dtype)
    zp, wp = solve_plain(a0)
    assert torch.equal(w1, wp) and torch.equal(a1, zp)      # the fallback is the plain solver
    m = 11
    am = a0.clone()
    am, wm, _, _ = eg.eigen_batch_warm(am, eye, max_sweeps=30, m=m)
    torch.cuda.synchronize()
    assert torch.equal(wm, w) and torch.equal(am[:, :m], a[:, :m])
