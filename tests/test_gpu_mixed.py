"""GPU: the mixed-precision solver of the FP64 interface (solver MIXED: FP32
solve, FP64 refinement on the matrix cores, plain FP64 fallback) against the
plain FP64 solve and the CPU oracle.

Bar per matrix (FP64 interface):
  * relative residual ||A Z - Z D||_F / max|w| <= 1e-12 and orthogonality
    ||Z^T Z - I||_F <= n TOL (the reference checker's threshold, eigen_GPU_check.cu:71-74),
    both from the GPU checker in FP64;
  * eigenvalues ascending and within TOL max|w| of the plain FP64 solve and of
    the oracle (on a sample).
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import workloads as wl

pytestmark = pytest.mark.gpu
DEV = "cuda"
RESID = 1e-12


def ref_tol(n):
    return 2.0 ** -52 * 1023.0 * np.sqrt(n + 16.0)


def full_sym(a0):
    """(L, n, nm) upper-triangle batch -> full symmetric device tensor of the same layout."""
    n = a0.shape[1]
    up = torch.triu(a0[:, :, :n].transpose(1, 2))          # row-major view, upper triangle
    full = up + up.transpose(1, 2) - torch.diag_embed(torch.diagonal(up, dim1=1, dim2=2))
    out = a0.clone()
    out[:, :, :n] = full.transpose(1, 2)
    return out


def solve(a0, solver):
    a = a0.clone()
    a, w, info = eg.eigen_batch(a, solver=solver)
    torch.cuda.synchronize()
    return a, w, info


def check_mixed(a0_np, iters_bits=0, oracle_sample=4):
    L, n, nm = a0_np.shape
    a0 = torch.from_numpy(np.nan_to_num(a0_np)).to(DEV)
    z, w, info = solve(a0, eg.MIXED | iters_bits)
    zp, wp, infop = solve(a0, eg.IMTQL2)
    assert int(info.abs().max()) == 0, f"info {info.cpu().numpy()[:8]}"
    out = eg.check(full_sym(a0), w, z)
    wmax = w.abs().amax(dim=1).double().clamp_min(1e-300)
    rel = (out[:, 0] / wmax).cpu().numpy()
    orth = out[:, 1].cpu().numpy()
    tol = ref_tol(n)
    assert rel.max() <= RESID, f"relative residual {rel.max():.3e} > {RESID}"
    assert orth.max() <= n * tol, f"orthogonality {orth.max():.3e} > {n * tol:.3e}"
    wn = w.cpu().numpy()
    assert np.all(np.diff(wn, axis=1) >= 0)
    scale = np.abs(wn).max(axis=1, keepdims=True)
    assert np.max(np.abs(wn - wp.cpu().numpy()) / scale) <= tol
    if oracle_sample:
        idx = np.linspace(0, L - 1, min(L, oracle_sample)).astype(int)
        _, wo, io = oracle.eig_batch(np.ascontiguousarray(np.nan_to_num(a0_np[idx])), oracle.IMTQL2)
        assert io.max() == 0
        assert np.max(np.abs(wn[idx] - wo) / scale[idx]) <= tol
    return rel, orth


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 100, 129, 256])
def test_mixed_sym_rand(n):
    print(f"[mixed] sym_rand n={n}", flush=True)
    nm = n + (n & 1)
    a0 = wl.sym_rand(9, n, nm, dtype=np.float64, seed=300 + n)
    check_mixed(a0)


@pytest.mark.parametrize("n", [64, 256])
def test_mixed_clustered_qdq(n):
    """The mixed-precision workload's family: A = Q diag(lam) Q^T with lam ~ U[-1,1]
    and 5 % of the eigenvalues in pairs split by 1e-8 (clusters below FP32
    resolution, resolved by the refinement's Rayleigh-Ritz step)."""
    print(f"[mixed] clustered qdq n={n}", flush=True)
    L = 24 if n == 256 else 64
    lam = wl.clustered_uniform(L, n, frac=0.05, gap=1e-8, seed=77 + n)
    a0 = wl.qdq(lam, dtype=np.float64, seed=78 + n)
    rel, _ = check_mixed(a0)
    print(f"[mixed] n={n} max relative residual {rel.max():.3e}", flush=True)


@pytest.mark.parametrize("n", [65, 100, 200, 256])
def test_mixed_with_the_fp16_blocked_back_transform(n):
    """The FP32 stage with the back-transformation as blocked reflectors on the
    FP16 matrix cores (EIGEN_GPU_BATCH_MIXED_WY16): the same bar, clustered
    spectrum, enough matrices for the large-order split path."""
    print(f"[mixed] fp16 blocked back-transform n={n}", flush=True)
    L = 520
    lam = wl.clustered_uniform(L, n, frac=0.05, gap=1e-8, seed=90 + n)
    a0 = wl.qdq(lam, dtype=np.float64, seed=91 + n)
    check_mixed(a0, iters_bits=eg.FLAG_MIXED_WY16, oracle_sample=2)


def test_mixed_large_clusters_take_the_fallback():
    """A triple eigenvalue (resolved by Rayleigh-Ritz), a 12-fold one and the
    identity (clusters beyond the Rayleigh-Ritz size: the plain FP64 fallback)."""
    print("[mixed] large clusters / fallback", flush=True)
    n = 40
    lam = np.stack([
        np.r_[np.full(3, 0.5), np.linspace(-1, 1, n - 3)],
        np.r_[np.full(12, -0.25), np.linspace(0.1, 1, n - 12)],
        np.ones(n),
        np.linspace(-1, 1, n),
    ])
    a0 = wl.qdq(lam, dtype=np.float64, seed=5)
    check_mixed(a0, oracle_sample=0)


def test_mixed_one_iteration_falls_back_and_stays_correct():
    """One refinement iteration leaves the measures above the acceptance bound,
    so every matrix goes to the plain FP64 solver: results still meet the bar."""
    print("[mixed] one iteration", flush=True)
    a0 = wl.sym_rand(6, 48, 48, dtype=np.float64, seed=9)
    check_mixed(a0, iters_bits=1 << eg.MIX_ITERS_SHIFT)


def test_mixed_chunks_and_padding():
    """More matrices than one refinement chunk (2048), nm > n with NaN padding
    rows left untouched."""
    print("[mixed] chunks", flush=True)
    n, nm, L = 12, 13, 2100
    a0_np = wl.sym_rand(L, n, nm, dtype=np.float64, seed=11)
    a0 = torch.from_numpy(a0_np).to(DEV)
    z, w, info = solve(a0, eg.MIXED)
    assert int(info.abs().max()) == 0
    assert bool(torch.isnan(z[:, :, n:]).all()), "padding rows were written"
    out = eg.check(full_sym(torch.nan_to_num(a0)), w, torch.nan_to_num(z))
    rel = (out[:, 0] / w.abs().amax(dim=1)).max().item()
    assert rel <= RESID


def test_mixed_rejected_on_fp32_and_large_orders():
    a = torch.zeros((2, 8, 8), dtype=torch.float32, device=DEV)
    with pytest.raises(RuntimeError):
        eg.eigen_batch(a, solver=eg.MIXED, workspace=torch.empty(1024, dtype=torch.uint8, device=DEV))
    assert eg.bufsize(4, 300, 300, 300, 8, eg.MIXED) == 0
