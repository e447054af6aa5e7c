"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar (per matrix, per the reference's own checker, eigen_GPU_check.cu:34-81):
  * eigenvalues: |w_gpu - w_oracle| <= TOL * max(|w_1|, |w_n|), with
    TOL = eps*(2*512-1 | 2*16-1)*sqrt(n+16) (the reference's pass threshold);
  * residual ||A Z - Z D||_F <= max|w| * TOL and orthogonality ||Z^T Z - I||_F
    <= n * TOL for the GPU eigenpairs (both checked by the oracle's checker);
  * eigenvectors of well separated eigenvalues agree with the oracle's up to
    sign: 1 - |<z_gpu, z_oracle>| <= 8 * n * eps / relgap.
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


def run_gpu(a_np, solver=eg.IMTQL2, abi=False, path="auto"):
    t = torch.from_numpy(np.ascontiguousarray(a_np)).to(DEV)
    t, w, info = eg.eigen_batch(t, solver=solver, abi=abi, path=path)
    torch.cuda.synchronize()
    return (t.cpu().numpy(), w.cpu().numpy(),
            None if info is None else info.cpu().numpy())


def check_batch(a0, z, w, dtype, wo=None, zo=None, vec_check=True):
    L, n, nm = a0.shape
    tol = ref_tol(n, dtype)
    full = wl.unpack(a0)
    for k in range(L):
        wk = w[k]
        scale = max(abs(wk[0]), abs(wk[-1]), np.finfo(dtype).tiny)
        assert np.all(np.diff(wk) >= 0), f"matrix {k}: eigenvalues not ascending"
        if wo is not None:
            err = np.max(np.abs(wk - wo[k])) / scale
            assert err <= tol, f"matrix {k}: eigenvalue error {err:.3e} > {tol:.3e}"
        a_full = np.zeros((n, nm), dtype=dtype)
        a_full[:, :n] = full[k].T
        out = oracle.check(a_full, wk, z[k])
        assert int(out[3]) == 0, (f"matrix {k}: checker flags {int(out[3])} "
                                  f"err1={out[0]:.3e} err2={out[1]:.3e} TOL={out[2]:.3e}")
        if vec_check and zo is not None and n > 1:
            eps = np.finfo(dtype).eps
            gaps = np.full(n, np.inf)
            dw = np.diff(wo[k])
            gaps[:-1] = np.minimum(gaps[:-1], dw)
            gaps[1:] = np.minimum(gaps[1:], dw)
            rel = gaps / scale
            ok = rel > 1e3 * np.sqrt(eps)
            dots = np.abs(np.einsum("jr,jr->j", z[k][:, :n].astype(np.float64),
                                    zo[k][:, :n].astype(np.float64)))
            bound = 8.0 * n * eps / np.maximum(rel, 1e-300)
            bad = ok & (1.0 - dots > bound)
            assert not bad.any(), f"matrix {k}: eigenvector mismatch at {np.nonzero(bad)[0][:5]}"


SIZES = [1, 2, 3, 4, 5, 7, 8, 9, 12, 15, 16, 17, 24, 31, 32, 33, 47, 48, 49, 63, 64,
         65, 66, 96, 100, 127, 128]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", SIZES)
def test_sym_rand_vs_oracle(n, dtype):
    print(f"[parity] sym_rand n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    L = 37
    a0 = wl.sym_rand(L, n, nm, dtype=dtype, seed=1000 + n)
    zo, wo, info_o = oracle.eig_batch(np.nan_to_num(a0), oracle.IMTQL2)
    assert info_o.max() == 0
    z, w, info = run_gpu(a0)
    assert info.max() == 0
    # padding rows must be left untouched (NaN in, NaN out)
    if nm > n:
        assert np.isnan(z[:, :, n:]).all()
    check_batch(np.nan_to_num(a0), z, w, dtype, wo, zo)


SPLIT_SIZES = [1, 2, 3, 5, 8, 9, 16, 17, 31, 32, 33, 48, 63, 64]


@pytest.mark.parametrize("path", ["split", "split-fallback", "split-lds"])
@pytest.mark.parametrize("solver", [eg.IMTQL2, eg.TQL2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", SPLIT_SIZES)
def test_split_path_vs_oracle(n, dtype, solver, path):
    print(f"[parity] {path} solver={solver} n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    L = 70
    a0 = wl.sym_rand(L, n, nm, dtype=dtype, seed=5000 + n)
    zo, wo, info_o = oracle.eig_batch(np.nan_to_num(a0), oracle.TQL2 if solver == eg.TQL2 else oracle.IMTQL2)
    assert info_o.max() == 0
    z, w, info = run_gpu(a0, solver=solver, path=path)
    assert info.max() == 0 and info.min() == 0
    if nm > n:
        assert np.isnan(z[:, :, n:]).all()
    check_batch(np.nan_to_num(a0), z, w, dtype, wo, zo)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [2, 7, 16, 31, 48, 64])
def test_ql_register_and_lds_kernels_agree(n, dtype):
    """The register lockstep QL kernel and the LDS lane kernel run the same
    per-step arithmetic: eigenvalues and eigenvectors agree to a few ulps
    (bit-identical unless the compiler contracts an expression differently)."""
    print(f"[parity] ql reg vs lds n={n} dtype={np.dtype(dtype).name}", flush=True)
    a0, _ = wl.letkf_y(150, k=n, p=3 * n + 10, rho=1.05, sigma=0.1, dtype=dtype, seed=100 + n)
    z1, w1, i1 = run_gpu(a0, path="split")
    z2, w2, i2 = run_gpu(a0, path="split-lds")
    assert (i1 == 0).all() and (i2 == 0).all()
    eps = np.finfo(dtype).eps
    scale = np.abs(w2).max(axis=1, keepdims=True)
    assert np.max(np.abs(w1 - w2) / scale) <= 16 * eps
    print(f"  bitwise equal: w {np.array_equal(w1, w2)} z {np.array_equal(z1, z2)}", flush=True)
    check_batch(a0, z1, w1, dtype, w2, z2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_split_and_fused_agree_on_letkf(dtype):
    a0, _ = wl.letkf_y(300, k=48, p=300, rho=1.05, sigma=0.1, dtype=dtype, seed=17)
    z1, w1, _ = run_gpu(a0, path="fused")
    z2, w2, _ = run_gpu(a0, path="split")
    tol = ref_tol(48, dtype)
    assert np.max(np.abs(w1 - w2) / np.abs(w1).max(axis=1, keepdims=True)) <= tol
    check_batch(a0, z2, w2, dtype, w1, z1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [129, 160, 255, 256, 300, 512])
def test_large_n_vs_oracle(n, dtype):
    print(f"[parity] large sym_rand n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    L = 5 if n <= 256 else 2
    a0 = wl.sym_rand(L, n, nm, dtype=dtype, seed=3000 + n)
    zo, wo, info_o = oracle.eig_batch(np.nan_to_num(a0), oracle.IMTQL2)
    assert info_o.max() == 0
    z, w, info = run_gpu(a0)
    assert info.max() == 0
    check_batch(np.nan_to_num(a0), z, w, dtype, wo, zo)


LARGE_SPLIT_SIZES = [65, 96, 129, 160, 200, 255, 256, 257, 300, 384, 512]


@pytest.mark.parametrize("path", ["split", "split-fallback", "split-full", "split-ws", "split-reg128",
                                  "split-reg128-fallback"])
@pytest.mark.parametrize("solver", [eg.IMTQL2, eg.TQL2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", LARGE_SPLIT_SIZES)
def test_large_split_vs_oracle(n, dtype, solver, path):
    """Orders beyond 64 through the split path (workgroup tridiagonalisation,
    lane-per-matrix QL with rotation log, row-block apply), through its
    overflow fallback (log capacity forced to 1), for n <= 256 with the
    full-storage and the workspace-resident lower-triangle tridiagonalisations
    instead of the register-resident one, and, for
    FP32 n <= 128, through the register split path (two-wave
    tridiagonalisation, 64-row-block apply) and its fallback."""
    if path in ("split-full", "split-ws") and n > 256:
        pytest.skip("the full-storage form is the only one beyond n = 256")
    if path.startswith("split-reg128") and (dtype != np.float32 or n > 128):
        pytest.skip("the register split path serves FP32 64 < n <= 128")
    print(f"[parity] large {path} solver={solver} n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    L = 7 if n <= 256 else 3
    a0 = wl.sym_rand(L, n, nm, dtype=dtype, seed=7000 + n)
    zo, wo, info_o = oracle.eig_batch(np.nan_to_num(a0), oracle.TQL2 if solver == eg.TQL2 else oracle.IMTQL2)
    assert info_o.max() == 0
    z, w, info = run_gpu(a0, solver=solver, path=path)
    assert info.max() == 0 and info.min() == 0
    if nm > n:
        assert np.isnan(z[:, :, n:]).all()
    check_batch(np.nan_to_num(a0), z, w, dtype, wo, zo)


@pytest.mark.parametrize("path", ["split", "split-full", "split-ws", "fused", "split-reg128"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [65, 128, 129, 256, 300])
def test_large_upper_triangle_only(n, dtype, path):
    """Only the upper triangle (with the diagonal) is read: NaN in the strictly
    lower triangle changes nothing, bit for bit."""
    if path == "split-reg128" and (dtype != np.float32 or n > 128):
        pytest.skip("the register split path serves FP32 64 < n <= 128")
    print(f"[parity] upper-only {path} n={n} dtype={np.dtype(dtype).name}", flush=True)
    L = 5
    a0 = wl.sym_rand(L, n, n, dtype=dtype, seed=7100 + n)
    a1 = a0.copy()
    r, c = np.tril_indices(n, -1)              # A(r, c), r > c  <->  a[k, c, r]
    a1[:, c, r] = np.nan
    z0, w0, i0 = run_gpu(a0, path=path)
    z1, w1, i1 = run_gpu(a1, path=path)
    assert i0.max() == 0 and i1.max() == 0
    np.testing.assert_array_equal(w1, w0)
    np.testing.assert_array_equal(z1, z0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [97, 200, 256])
def test_register_tridiagonalisation_many_matrices_per_workgroup(n, dtype):
    """The register-resident tridiagonalisation with a grid of 128 workgroups
    (grid code 1 in bits 20-23), so each workgroup reduces several matrices in
    turn; eigenpairs against the oracle on a sample and the GPU checker on all,
    and eigenvalues against the workspace-resident kernel."""
    print(f"[parity] register tri, persistent grid n={n} dtype={np.dtype(dtype).name}", flush=True)
    L = 300
    a0 = wl.sym_rand(L, n, n, dtype=dtype, seed=7300 + n)
    t = torch.from_numpy(a0).to(DEV)
    z, w, info = eg.eigen_batch(t.clone(), solver=eg.IMTQL2 | eg.FLAG_SPLIT | (1 << 20))
    z2, w2, info2 = eg.eigen_batch(t.clone(), path="split-ws")
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0 and int(info2.abs().max()) == 0
    out = eg.check(t, w, z).cpu().numpy()
    assert int(out[:, 3].max()) == 0, f"checker flags on {int((out[:, 3] != 0).sum())} matrices"
    wn, w2n = w.cpu().numpy(), w2.cpu().numpy()
    scale = np.abs(wn).max(axis=1, keepdims=True)
    assert np.max(np.abs(wn - w2n) / scale) <= ref_tol(n, dtype)
    idx = [0, 1, 150, L - 1]
    zo, wo, _ = oracle.eig_batch(np.ascontiguousarray(a0[idx]), oracle.IMTQL2)
    check_batch(a0[idx], z.cpu().numpy()[idx], wn[idx], dtype, wo, zo)


@pytest.mark.parametrize("dtype", [np.float64])
def test_large_split_letkf_and_qdq(dtype):
    """n = 256 clustered spectra (LETKF-like, and pairs split by 1e-8) through
    the large split path."""
    qdq = wl.qdq(wl.clustered_uniform(6, 256, frac=0.05, gap=1e-8, seed=6), dtype=dtype, seed=6)
    for a0 in (wl.letkf_y(6, k=256, p=400, rho=1.05, sigma=0.1, dtype=dtype, seed=5)[0], qdq):
        zo, wo, _ = oracle.eig_batch(a0, oracle.IMTQL2)
        z, w, info = run_gpu(a0, path="split")
        assert info.max() == 0
        check_batch(a0, z, w, dtype, wo, zo)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [5, 32, 64, 100, 200])
def test_tql2_variant(n, dtype):
    print(f"[parity] tql2 n={n} dtype={np.dtype(dtype).name}", flush=True)
    a0 = wl.sym_rand(19, n, n, dtype=dtype, seed=77 + n)
    zo, wo, _ = oracle.eig_batch(a0, oracle.TQL2)
    z, w, info = run_gpu(a0, solver=eg.TQL2)
    assert info.max() == 0
    check_batch(a0, z, w, dtype, wo, zo)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 9, 16, 20, 32, 33, 48, 64])
def test_jacobi_vs_oracle(n, dtype):
    print(f"[parity] jacobi n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    a0 = wl.sym_rand(21, n, nm, dtype=dtype, seed=4000 + n)
    zo, wo, _ = oracle.eig_batch(np.nan_to_num(a0), oracle.JACOBI)
    z, w, info = run_gpu(a0, solver=eg.JACOBI)
    assert info.min() == 0 and info.max() == 0
    if nm > n:
        assert np.isnan(z[:, :, n:]).all()
    check_batch(np.nan_to_num(a0), z, w, dtype, wo, zo)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_jacobi_closed_forms(dtype):
    n = 40
    a0, lam = wl.letkf_helmert(6, n, n, dtype=dtype, seed=8)
    z, w, info = run_gpu(a0, solver=eg.JACOBI)
    assert info.max() == 0
    assert np.max(np.abs(w - lam) / lam[:, -1:]) <= ref_tol(n, dtype)
    check_batch(a0, z, w, dtype, vec_check=False)
    a1 = wl.frank(3, 30, 30, dtype=dtype)
    z, w, info = run_gpu(a1, solver=eg.JACOBI)
    assert np.max(np.abs(w - wl.frank_eigenvalues(30)) / w[:, -1:]) <= ref_tol(30, dtype)
    check_batch(a1, z, w, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [3, 10, 32, 48, 64, 97, 128, 256])
def test_frank_closed_form(n, dtype):
    print(f"[parity] frank n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    a0 = wl.frank(8, n, nm, dtype=dtype)
    z, w, info = run_gpu(a0)
    lam = wl.frank_eigenvalues(n)
    tol = ref_tol(n, dtype)
    assert np.max(np.abs(w - lam[None, :])) <= tol * lam[-1]
    check_batch(np.nan_to_num(a0), z, w, dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_letkf_helmert_closed_form(dtype):
    n = 48
    a0, lam = wl.letkf_helmert(16, n, n, dtype=dtype, seed=5)
    z, w, info = run_gpu(a0)
    tol = ref_tol(n, dtype)
    assert np.max(np.abs(w - lam) / lam[:, -1:]) <= tol
    check_batch(a0, z, w, dtype, vec_check=False)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_abi_entry_matches_extended(dtype):
    a0 = wl.sym_rand(50, 20, 20, dtype=dtype, seed=3)
    z1, w1, _ = run_gpu(a0)
    z2, w2, _ = run_gpu(a0, abi=True)
    assert np.array_equal(w1, w2) and np.array_equal(z1, z2)


def test_special_matrices():
    for dtype in (np.float32, np.float64):
        n = 16
        # identity (multiplicity n), zero matrix, diagonal with repeats
        eye = np.broadcast_to(np.eye(n), (1, n, n))
        zero = np.zeros((1, n, n))
        diag = np.broadcast_to(np.diag(np.repeat([3.0, -1.0, 2.0, 2.0], 4)), (1, n, n))
        a0 = np.concatenate([eye, zero, diag]).astype(dtype)
        z, w, info = run_gpu(a0)
        assert info.max() == 0
        assert np.allclose(w[0], 1.0) and np.allclose(w[1], 0.0)
        assert np.allclose(w[2], np.sort(np.repeat([3.0, -1.0, 2.0, 2.0], 4)))
        for k in range(3):
            zz = z[k].astype(np.float64)
            assert np.allclose(zz @ zz.T, np.eye(n), atol=1e-5)


def test_invalid_arguments_are_noops():
    a = torch.ones((2, 4, 4), dtype=torch.float32, device=DEV)
    w = torch.zeros((2, 4), dtype=torch.float32, device=DEV)
    ws = torch.empty(256, dtype=torch.uint8, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    lib = eg.lib()
    for (L, nm, n, m) in [(0, 4, 4, 4), (2, 3, 4, 4), (2, 4, 0, 1), (2, 4, 4, 0), (2, 4, 4, 5)]:
        lib.eigen_GPU_batch_FP(L, nm, n, m, a.data_ptr(), w.data_ptr(), ws.data_ptr(), s)
        rc = lib.eigen_GPU_batch_FP_x(L, nm, n, m, a.data_ptr(), w.data_ptr(), ws.data_ptr(), s, 0, 0)
        assert rc != 0
    torch.cuda.synchronize()
    assert torch.all(a == 1) and torch.all(w == 0)


@pytest.mark.parametrize("n,dtype,L", [(48, np.float32, 2500), (200, np.float64, 600)])
def test_autotune_picks_a_candidate_and_stays_correct(n, dtype, L):
    """The autotuner times its candidates on the batch, returns one of them,
    leaves a_src untouched and the chosen configuration's eigenpairs correct;
    a later default call of the same shape still solves correctly."""
    print(f"[tune] n={n} dtype={np.dtype(dtype).name} L={L}", flush=True)
    eg.tune_reset()
    a0 = wl.sym_rand(L, n, n, dtype=dtype, seed=900 + n)
    src = torch.from_numpy(a0.copy()).to(DEV)
    best, best_ms, table, z, w = eg.autotune(src)
    torch.cuda.synchronize()
    assert best in table and best_ms == pytest.approx(min(table.values()))
    assert len(table) >= 2
    assert torch.equal(src.cpu(), torch.from_numpy(a0))
    full = torch.from_numpy(np.ascontiguousarray(np.transpose(wl.unpack(a0), (0, 2, 1)))).to(DEV)
    out = eg.check(full, w, z).cpu().numpy()
    assert (out[:, 3] == 0).all()
    z2, w2, info = run_gpu(a0)
    assert info.max() == 0
    out2 = eg.check(full, torch.from_numpy(w2).to(DEV), torch.from_numpy(z2).to(DEV)).cpu().numpy()
    assert (out2[:, 3] == 0).all()
    eg.tune_reset()
