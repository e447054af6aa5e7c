"""GPU: the output options of the _x entry points on every path.

* m < n (eigen_GPU_batch.cu:89,133 pass m through): all n eigenvalues, the
  eigenvectors of the first m written to columns 0..m-1 -- bitwise the same
  values as the m = n call of the same path;
* unsorted (the reference's DO_SORT 0, eigen_GPU_batch.cu:5,10): the same
  eigenpairs in the QL / Jacobi order, i.e. sorting them gives back the sorted
  call's eigenvalues bitwise and the eigenvector columns match, and every pair
  passes the reference checker.
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import workloads as wl

pytestmark = pytest.mark.gpu
DEV = "cuda"

# (label, dtype, n, L, solver, path)
CASES = [
    ("fused-32", np.float64, 32, 40, eg.IMTQL2, "fused"),
    ("split-48", np.float32, 48, 300, eg.IMTQL2, "split"),
    ("split-lds-tql2", np.float32, 24, 200, eg.TQL2, "split-lds"),
    ("split-fallback", np.float32, 20, 100, eg.IMTQL2, "split-fallback"),
    ("reg128-100", np.float32, 100, 70, eg.IMTQL2, "split-reg128"),
    ("large-split-130", np.float64, 130, 40, eg.IMTQL2, "split"),
    ("workgroup-150", np.float64, 150, 6, eg.IMTQL2, "fused"),
    ("jacobi-30", np.float64, 30, 50, eg.JACOBI, "auto"),
    ("mixed-70", np.float64, 70, 20, eg.MIXED, "auto"),
]


def run(a0, solver, path, m=None, unsorted=False):
    a = a0.clone()
    a, w, info = eg.eigen_batch(a, m=m, solver=solver, path=path, unsorted=unsorted)
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0
    return a.cpu().numpy(), w.cpu().numpy()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_partial_eigenvectors(case):
    label, dtype, n, L, solver, path = case
    print(f"[api] m<n {label}", flush=True)
    a0 = torch.from_numpy(wl.sym_rand(L, n, n, dtype=dtype, seed=n)).to(DEV)
    z, w = run(a0, solver, path)
    m = max(1, n // 3)
    zm, wm = run(a0, solver, path, m=m)
    assert np.array_equal(w, wm)
    assert np.array_equal(z[:, :m, :], zm[:, :m, :])


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_unsorted_output(case):
    label, dtype, n, L, solver, path = case
    print(f"[api] unsorted {label}", flush=True)
    a0n = wl.sym_rand(L, n, n, dtype=dtype, seed=n + 1)
    a0 = torch.from_numpy(a0n).to(DEV)
    z, w = run(a0, solver, path)
    zu, wu = run(a0, solver, path, unsorted=True)
    # the same multiset of eigenvalues (the sorted call uses the reference's
    # selection sort: a stable order of the same values)
    assert np.array_equal(np.sort(wu, axis=1, kind="stable"), w)
    full = wl.unpack(a0n)
    for k in range(0, L, max(1, L // 8)):
        a_full = np.ascontiguousarray(full[k].T)
        out = oracle.check(a_full, wu[k], zu[k])
        assert int(out[3]) == 0, (label, k, out[:3])
        perm = np.argsort(wu[k], kind="stable")
        # columns agree up to sign where eigenvalues are simple
        dots = np.abs(np.einsum("jr,jr->j", zu[k][perm].astype(np.float64), z[k].astype(np.float64)))
        gaps = np.minimum(np.r_[np.inf, np.diff(w[k])], np.r_[np.diff(w[k]), np.inf])
        ok = gaps > 1e-3 * np.abs(w[k]).max()
        assert np.all(dots[ok] > 1 - 1e-3), (label, k)
    if solver != eg.MIXED and n > 8:
        # the QL order is not the sorted one for these matrices
        assert not np.array_equal(wu, w)


@pytest.mark.parametrize("dtype,n,nm,L,path", [(np.float32, 48, 49, 3000, "auto"), (np.float64, 33, 34, 200, "auto"),
                                               (np.float32, 100, 100, 50, "auto"), (np.float64, 20, 20, 7, "fused")])
def test_interleaved_layout(dtype, n, nm, L, path):
    """The interleaved layout (matrix index fastest) gives the strided call's
    results bitwise, transposed."""
    print(f"[api] interleaved n={n} nm={nm} L={L} {np.dtype(dtype).name}", flush=True)
    a0 = torch.from_numpy(wl.sym_rand(L, n, nm, dtype=dtype, seed=3 * n)).to(DEV)
    z, w = run(a0, eg.IMTQL2, path)
    ai = a0.permute(1, 2, 0).contiguous()             # (n, nm, L): element (r, c, k) at k + L (r + c nm)
    ai, wi, info = eg.eigen_batch_interleaved(ai, L, nm, n, solver=eg.IMTQL2 | eg.PATHS[path])
    torch.cuda.synchronize()
    assert int(info.abs().max()) == 0
    zi = ai.permute(2, 0, 1).cpu().numpy()
    assert np.array_equal(wi.t().cpu().numpy(), w)
    assert np.array_equal(zi[:, :, :n], z[:, :, :n])
    assert np.isnan(zi[:, :, n:]).all()                # padding rows carried through untouched


@pytest.mark.parametrize("dtype,n,L,solver,shift", [(np.float32, 64, 5000, eg.IMTQL2, 3),
                                                    (np.float32, 100, 4100, eg.IMTQL2, 2),
                                                    (np.float64, 100, 1200, eg.IMTQL2, 3),
                                                    (np.float64, 90, 600, eg.MIXED, 2)])
def test_smaller_chunks_give_the_same_results(dtype, n, L, solver, shift):
    """Bits 18-19 (the batch chunk divided by 2^shift): bitwise the same
    eigenpairs as the default chunk, in a smaller workspace."""
    print(f"[api] chunk/{1 << shift} n={n} L={L} {np.dtype(dtype).name} solver={solver}", flush=True)
    a0 = torch.from_numpy(wl.sym_rand(L, n, n, dtype=dtype, seed=n + shift)).to(DEV)
    path = "split" if (dtype == np.float64 and solver == eg.IMTQL2) else "auto"
    z, w = run(a0, solver, path)
    zc, wc = run(a0, solver | (shift << eg.CHUNK_SHIFT), path)
    assert np.array_equal(w, wc) and np.array_equal(z, zc)
