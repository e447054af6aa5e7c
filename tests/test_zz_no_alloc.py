"""The no-allocation entry points eigen_GPU_batch_FP_noalloc / _DP_noalloc
against the workspace entry points eigen_GPU_batch_FP_x / _DP_x whose kernels
they share.

Two calls in a row on each element type, the second with a smaller batch:
* when the whole batch fits in the library's fixed scratch array, the path and
  the kernels are the _x call's, so the results are bitwise identical;
* when it does not (FP64 n = 200, L = 600), the batch is solved in sub-batches
  that take the workgroup kernel instead of the large split path: eigenvalues
  agree within the reference checker's TOL * max|w| (eigen_GPU_check.cu:71)
  and every eigenpair passes its residual and orthogonality tests.
"""
import ctypes

import numpy as np
import pytest
import torch

import eigeng_batched as eg
import workloads as wl

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _lib():
    L = eg.lib()
    for nme in ("eigen_GPU_batch_FP_noalloc", "eigen_GPU_batch_DP_noalloc"):
        f = getattr(L, nme)
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
        f.restype = ctypes.c_int
    L.eigen_GPU_batch_noalloc_scratch_bytes.restype = ctypes.c_size_t
    return L


def noalloc(a_np, solver=eg.IMTQL2):
    t = torch.from_numpy(np.ascontiguousarray(a_np)).to(DEV)
    L, n, nm = t.shape
    w = torch.empty((L, n), dtype=t.dtype, device=DEV)
    info = torch.full((L,), -7, dtype=torch.int32, device=DEV)
    f = _lib().eigen_GPU_batch_FP_noalloc if t.dtype == torch.float32 else _lib().eigen_GPU_batch_DP_noalloc
    rc = f(L, nm, n, n, t.data_ptr(), w.data_ptr(), torch.cuda.current_stream().cuda_stream, solver,
           info.data_ptr())
    torch.cuda.synchronize()
    return rc, t.cpu().numpy(), w.cpu().numpy(), info.cpu().numpy()


def with_workspace(a_np, solver=eg.IMTQL2):
    t = torch.from_numpy(np.ascontiguousarray(a_np)).to(DEV)
    t, w, info = eg.eigen_batch(t, solver=solver)
    torch.cuda.synchronize()
    return t.cpu().numpy(), w.cpu().numpy(), info.cpu().numpy()


def ref_tol(n, dtype):
    eps = float(np.finfo(dtype).eps)
    return eps * (1023.0 if dtype == np.float64 else 31.0) * np.sqrt(n + 16.0)


def test_scratch_size_is_stated():
    assert _lib().eigen_GPU_batch_noalloc_scratch_bytes() == 256 << 20


@pytest.mark.parametrize("shape", [(3000, 64), (100, 20)])
def test_fp32_two_calls_match_the_workspace_entry_bitwise(shape):
    """First (3000, 64): the split path, about 200 MB of workspace, one
    sub-batch; then (100, 20): the fused kernel."""
    L, n = shape
    print(f"[noalloc] FP32 L={L} n={n}", flush=True)
    a0 = wl.sym_rand(L, n, n + (n & 1), dtype=np.float32, seed=31 + n)
    assert eg.bufsize(L, n + (n & 1), n, n, 4) <= _lib().eigen_GPU_batch_noalloc_scratch_bytes()
    rc, z1, w1, i1 = noalloc(a0)
    assert rc == 0
    z2, w2, i2 = with_workspace(a0)
    assert (i1 == 0).all() and (i2 == 0).all()
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(z1, z2)


def test_fp64_sub_batches_then_a_smaller_call():
    """FP64 n = 200, L = 600: the large split path's workspace exceeds the
    scratch array, so the batch runs as sub-batches (checked to the reference
    checker's bar against the _x call); then a smaller call, n = 33, L = 40,
    which fits and matches bitwise."""
    from oracle import check
    L, n = 600, 200
    print(f"[noalloc] FP64 L={L} n={n} (sub-batches)", flush=True)
    assert eg.bufsize(L, n, n, n, 8) > _lib().eigen_GPU_batch_noalloc_scratch_bytes()
    a0 = wl.sym_rand(L, n, n, dtype=np.float64, seed=77)
    rc, z1, w1, i1 = noalloc(a0)
    assert rc == 0 and (i1 == 0).all()
    z2, w2, i2 = with_workspace(a0)
    assert (i2 == 0).all()
    tol = ref_tol(n, np.float64)
    scale = np.abs(w2).max(axis=1, keepdims=True)
    assert np.max(np.abs(w1 - w2) / scale) <= tol
    full = wl.unpack(a0)
    for k in range(0, L, 37):
        out = check(np.ascontiguousarray(full[k].T), w1[k], z1[k])
        assert int(out[3]) == 0, f"matrix {k}: {out}"
    print("[noalloc] FP64 L=40 n=33", flush=True)
    a1 = wl.sym_rand(40, 33, 34, dtype=np.float64, seed=78)
    rc, z1, w1, i1 = noalloc(a1)
    assert rc == 0 and (i1 == 0).all()
    z2, w2, _ = with_workspace(a1)
    np.testing.assert_array_equal(w1, w2)
    np.testing.assert_array_equal(z1, z2)


def test_rejections_launch_nothing():
    a = torch.ones((2, 4, 4), dtype=torch.float32, device=DEV)
    w = torch.zeros((2, 4), dtype=torch.float32, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    f = _lib().eigen_GPU_batch_FP_noalloc
    assert f(0, 4, 4, 4, a.data_ptr(), w.data_ptr(), s, 0, None) != 0          # bad L
    assert f(2, 4, 4, 4, a.data_ptr(), w.data_ptr(), s, 0x1000, None) != 0     # phase bits
    assert f(2, 4, 4, 4, a.data_ptr(), w.data_ptr(), s, 0x400, None) != 0      # test-only log bit
    assert f(2, 513, 513, 513, a.data_ptr(), w.data_ptr(), s, 0, None) != 0    # order beyond 512
    torch.cuda.synchronize()
    assert torch.all(a == 1) and torch.all(w == 0)
