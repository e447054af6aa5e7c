"""GPU parity on the exact input generators of the benchmark workloads, through
the paths those workloads take by default, against the CPU oracle.

Each FP32 case is also compared with an FP64 solve of the same (FP32) input:
FP32 eigenvalues must lie within the reference checker's FP32 bar
TOL * max|w| of the FP64 eigenvalues (TOL = eps (2*16-1) sqrt(n+16),
eigen_GPU_check.cu:71), and every GPU eigenpair must pass the checker's
residual and orthogonality tests (eigen_GPU_check.cu:34-81).

Workloads (generators in workloads.py; the reference's own families are at
main.cpp:63-195):
* n = 32 FP64, L = 1024, A = (B + B^T)/2, B ~ U[-1,1]: the fused register path;
* LETKF n = 64 FP32, A = (k-1)/rho I + Y^T Y, k = 64, p = 400, rho = 1.1,
  sigma = 0.1, L = 2048: the default split path (L >= 2048);
* uneven FP32 batch, orders U{8..128}, A = Q diag(lam) Q^T with lam
  log-spaced over 1..1e-6 in random order, with more than 4096 matrices of
  orders in (64, 128], so the split route runs for that class by count;
* n = 256 FP64, lam ~ U[-1,1] with 5 % of the eigenvalues in pairs split by
  1e-8: the large split path (L >= 512);
* LETKF n = 48 FP32 (k = 48, p = 300, rho = 1.05): the drifting-stream shape.
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


def solve(a_np, path="auto"):
    t = torch.from_numpy(np.ascontiguousarray(a_np)).to(DEV)
    t, w, info = eg.eigen_batch(t, path=path)
    torch.cuda.synchronize()
    return t.cpu().numpy(), w.cpu().numpy(), info.cpu().numpy()


def gpu_check_all(a0, z, w):
    """The GPU checker (reference metrics) over every matrix of the batch."""
    L, n, nm = a0.shape
    full = np.ascontiguousarray(np.transpose(wl.unpack(a0), (0, 2, 1)))
    out = eg.check(torch.from_numpy(full).to(DEV), torch.from_numpy(w).to(DEV),
                   torch.from_numpy(z).to(DEV)).cpu().numpy()
    bad = np.nonzero(out[:, 3])[0]
    assert bad.size == 0, f"{bad.size} matrices fail the checker, first {bad[:5]}: {out[bad[0]]}"
    return out


def compare_sample(a0, z, w, idx, dtype):
    """Oracle eigenvalues in type (and, for FP32, from the FP64 oracle on the
    same input) within TOL * max|w|; oracle checker on the GPU eigenpairs."""
    n = a0.shape[1]
    tol = ref_tol(n, dtype)
    sub = np.ascontiguousarray(a0[idx])
    _, wo, io = oracle.eig_batch(sub, oracle.IMTQL2)
    assert io.max() == 0
    scale = np.maximum(np.abs(wo[:, :1]), np.abs(wo[:, -1:]))
    err = np.max(np.abs(w[idx] - wo) / scale)
    assert err <= tol, f"eigenvalues vs the in-type oracle: {err:.3e} > {tol:.3e}"
    if dtype == np.float32:
        _, wd, iod = oracle.eig_batch(sub.astype(np.float64), oracle.IMTQL2)
        assert iod.max() == 0
        err64 = np.max(np.abs(w[idx].astype(np.float64) - wd) / np.maximum(np.abs(wd[:, :1]), np.abs(wd[:, -1:])))
        print(f"  FP32 eigenvalues vs FP64 oracle: {err64:.3e} (bar {tol:.3e})", flush=True)
        assert err64 <= tol, f"FP32 eigenvalues vs the FP64 oracle: {err64:.3e} > {tol:.3e}"
    full = wl.unpack(a0)
    for j, k in enumerate(idx):
        out = oracle.check(np.ascontiguousarray(full[k].T), w[k], z[k])
        assert int(out[3]) == 0, f"matrix {k}: {out}"


def test_random_symmetric_n32_fp64_L1024():
    print("[workload] n=32 FP64 L=1024 sym_rand", flush=True)
    a0 = wl.sym_rand(1024, 32, 32, dtype=np.float64, seed=0xE16E0001)
    assert eg.launch_params(1024, 32, 8)["numGR"] == 2          # fused: two 32-lane groups per wave
    z, w, info = solve(a0)
    assert info.max() == 0 and info.min() == 0
    gpu_check_all(a0, z, w)
    compare_sample(a0, z, w, np.arange(1024), np.float64)


def test_letkf_n64_fp32_split_path():
    print("[workload] LETKF n=64 FP32 k=64 p=400 rho=1.1 L=2048", flush=True)
    L = 2048
    a0, _ = wl.letkf_y(L, k=64, p=400, rho=1.1, sigma=0.1, dtype=np.float32, seed=0xE16E0002)
    z, w, info = solve(a0)                                     # default: the split path from 2048
    assert info.max() == 0 and info.min() == 0
    z2, w2, _ = solve(a0, path="split")
    np.testing.assert_array_equal(w, w2)                       # the default path is the split path
    np.testing.assert_array_equal(z, z2)
    gpu_check_all(a0, z, w)
    compare_sample(a0, z, w, np.linspace(0, L - 1, 256).astype(int), np.float32)


def test_letkf_n48_fp32_stream_shape():
    print("[workload] LETKF n=48 FP32 k=48 p=300 rho=1.05 L=4096", flush=True)
    L = 4096
    a0, _ = wl.letkf_y(L, k=48, p=300, rho=1.05, sigma=0.1, dtype=np.float32, seed=0xE16E0005)
    z, w, info = solve(a0)
    assert info.max() == 0
    gpu_check_all(a0, z, w)
    compare_sample(a0, z, w, np.linspace(0, L - 1, 256).astype(int), np.float32)


def test_clustered_pairs_n256_fp64_large_split():
    print("[workload] n=256 FP64 lam ~ U[-1,1], 5% pairs split by 1e-8, L=512", flush=True)
    L, n = 512, 256
    lam = wl.clustered_uniform(L, n, frac=0.05, gap=1e-8, seed=0xE16E0004)
    a0 = wl.qdq(lam, dtype=np.float64, seed=0xE16E0004)
    z, w, info = solve(a0)                                     # default: large split path (L >= 512)
    assert info.max() == 0
    gpu_check_all(a0, z, w)
    # eigenvalues against the prescribed spectrum (closed form) and the oracle
    tol = ref_tol(n, np.float64)
    lam_s = np.sort(lam, axis=1)
    assert np.max(np.abs(w - lam_s) / np.abs(lam_s).max(axis=1, keepdims=True)) <= tol
    compare_sample(a0, z, w, np.arange(0, L, 16), np.float64)


def test_uneven_logspaced_fp32_split_route_by_count():
    """Orders ~ U{8..128}; the (64, 128] class holds more than 4096 matrices, so
    the split route takes it by count (no forcing bit)."""
    rng = np.random.default_rng(0xE16E0003)
    count = 8400
    ns = rng.integers(8, 129, size=count)
    assert (ns > 64).sum() > 4096
    mats, lams = [], []
    for n in range(8, 129):
        idx = np.nonzero(ns == n)[0]
        if idx.size == 0:
            continue
        lam = wl.logspaced_shuffled(idx.size, n, 1.0, 1e-6, seed=int(n))
        A = wl.unpack(wl.qdq(lam, dtype=np.float64, seed=1000 + int(n)))
        for j, k in enumerate(idx):
            mats.append((k, A[j].astype(np.float32)))
    mats.sort(key=lambda t: t[0])
    print(f"[workload] uneven FP32 count={count}, {(ns > 64).sum()} in (64,128]", flush=True)
    a, nv = eg.pack_uneven([m for _, m in mats], device=DEV)
    a, w, info = eg.eigen_batch_uneven(a, nv, 128)
    torch.cuda.synchronize()
    assert int(info.max().item()) == 0 and int(info.min().item()) == 0
    res = eg.unpack_uneven(a, w, [int(x) for x in ns])
    tol_by_n = {}
    by_order = {}
    for k, (z, wk) in enumerate(res):
        by_order.setdefault(int(ns[k]), []).append(k)
    for n, ks in by_order.items():
        tol = ref_tol(n, np.float32)
        A = np.stack([mats[k][1] for k in ks])                 # (c, n, n), symmetric
        batch = np.ascontiguousarray(np.swapaxes(A, 1, 2))     # column-major view (c, n, n)
        _, wo, io = oracle.eig_batch(batch, oracle.IMTQL2, threads=0)
        _, wd, iod = oracle.eig_batch(batch.astype(np.float64), oracle.IMTQL2, threads=0)
        assert io.max() == 0 and iod.max() == 0
        wg = np.stack([res[k][1] for k in ks])
        scale = np.abs(wd).max(axis=1, keepdims=True)
        assert np.max(np.abs(wg - wo) / scale) <= tol, f"order {n}: vs the FP32 oracle"
        assert np.max(np.abs(wg.astype(np.float64) - wd) / scale) <= tol, f"order {n}: vs the FP64 oracle"
        for j in range(0, len(ks), 7):
            k = ks[j]
            z, wk = res[k]
            out = oracle.check(np.ascontiguousarray(A[j].T), wk, np.ascontiguousarray(z.T))
            assert int(out[3]) == 0, f"matrix {k} (order {n}): {out}"
        tol_by_n[n] = tol
    assert len(tol_by_n) == 121
