"""ctypes binding of libeigenGbatch (the C ABI in include/eigen_GPU_batch.h).

This is the binding a Python caller of the reference would add (the reference
ships only C/C++ headers, eigen_GPU_batch.h:26-30 and eigen_GPU_batch.hpp:104-157).
PyTorch is used only for device memory and the current HIP stream.  There is no
CPU fallback: every call goes through the HIP library, and loading fails loudly
when the library has not been built.

Tensor layout is the reference's batch-strided column-major layout viewed from
row-major torch: a tensor of shape (L, n, nm) where ``t[k, c, r]`` is element
(r, c) of matrix k (leading dimension nm >= n).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeigenGbatch.so")

IMTQL2, TQL2, JACOBI = 0, 1, 2
# FP64 interface only: FP32 solve + FP64 refinement on the matrix cores, with the
# plain FP64 solver as fallback; bits 28-29 set the refinement iterations (0: 3)
MIXED = 3
MIX_ITERS_SHIFT = 28
# with MIXED, 64 < n <= 256: the FP32 stage's back-transformation as blocked
# reflectors on the FP16 matrix cores (error-compensated split products)
FLAG_MIXED_WY16 = 0x8000000
# batched products (eigen_GPU_batch_gemm): matrix-core operand kinds
GEMM_F64, GEMM_F32, GEMM_F16, GEMM_BF16, GEMM_F16X3 = 0, 1, 2, 3, 4
# tuning flags OR-ed into the solver id of the _x entry points
FLAG_FUSED, FLAG_SPLIT, FLAG_TINYLOG = 0x100, 0x200, 0x400
FLAG_QL_LDS = 0x800
FLAG_TRI_FULL = 0x4000000
# bits 24-25: the large split path's tridiagonalisation kernel; 0 (default) the
# register-resident one for n <= 256, 1 the workspace-resident lower-triangle one
TRI_CFG_SHIFT = 24
FLAG_TRI_WS = 1 << TRI_CFG_SHIFT
FLAG_REG128 = 0x8000
# eigenpairs in the QL / Jacobi order instead of ascending (the reference's DO_SORT 0)
FLAG_UNSORTED = 0x40000000
# bits 18-19: the split paths' batch chunk divided by 2, 4 or 8 (smaller workspace, longer call)
CHUNK_SHIFT = 18
PATHS = {"auto": 0, "fused": FLAG_FUSED, "split": FLAG_SPLIT, "split-fallback": FLAG_SPLIT | FLAG_TINYLOG,
         "split-lds": FLAG_SPLIT | FLAG_QL_LDS, "split-full": FLAG_SPLIT | FLAG_TRI_FULL,
         "split-ws": FLAG_SPLIT | FLAG_TRI_WS,
         "split-reg128": FLAG_REG128, "split-reg128-fallback": FLAG_REG128 | FLAG_TINYLOG}
# split-path phases (a call with a subset runs only those kernels; the phases
# hand over through the workspace, so tri -> ql -> apply calls on one stream
# and one workspace compose the full solve, for batches of at most 16384)
PHASES = {"tri": 0x1000, "ql": 0x2000, "apply": 0x4000}
_lib = None


def build() -> None:
    """Compile the HIP library for gfx950 in-tree (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        for name in ("eigen_GPU_batch_FP", "eigen_GPU_batch_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp]
            f.restype = None
        for name in ("eigen_GPU_batch_FP_x", "eigen_GPU_batch_DP_x"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, ci, vp]
            f.restype = ci
        for name in ("eigen_GPU_batch_uneven_FP", "eigen_GPU_batch_uneven_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, vp, vp, vp, vp, vp, vp, vp, ci, vp]
            f.restype = ci
        for name in ("eigen_GPU_check_x_FP", "eigen_GPU_check_x_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp]
            f.restype = ci
        for name in ("eigen_GPU_check_FP", "eigen_GPU_check_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp]
            f.restype = None
        for name in ("eigen_GPU_batch_tune_FP", "eigen_GPU_batch_tune_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, vp, vp, vp, vp, vp, ci, vp, vp, ci, vp, vp, vp]
            f.restype = ci
        cf, cd = ctypes.c_float, ctypes.c_double
        for name, ct in (("FP", cf), ("DP", cd)):
            f = getattr(L, "eigen_GPU_batch_symdmt_" + name)
            f.argtypes = [ci, ci, ci, vp, vp, vp, vp]
            f.restype = ci
            f = getattr(L, "eigen_GPU_batch_matpow_" + name)
            f.argtypes = [ci, ci, ci, vp, vp, ct, ct, vp, vp]
            f.restype = ci
            f = getattr(L, "eigen_GPU_batch_gram_" + name)
            f.argtypes = [ci, ci, ci, ci, vp, ci, vp, ci, ct, vp, vp]
            f.restype = ci
        L.eigen_GPU_batch_tune_bufsize.argtypes = [ci, ci, ci, ci, ci]
        L.eigen_GPU_batch_tune_bufsize.restype = sz
        L.eigen_GPU_batch_tune_set.argtypes = [ci, ci, ci, ci, ci]
        L.eigen_GPU_batch_tune_set.restype = ci
        L.eigen_GPU_batch_launch_params.argtypes = [ci, ci, ci, ci, vp]
        L.eigen_GPU_batch_launch_params.restype = ci
        L.eigen_GPU_batch_tune_reset.argtypes = []
        L.eigen_GPU_batch_tune_reset.restype = None
        L.eigen_GPU_batch_uneven_bufsize.argtypes = [ci, ci, ci, ci]
        L.eigen_GPU_batch_uneven_bufsize.restype = sz
        L.eigen_GPU_batch_bufsize.argtypes = [ci, ci, ci, ci, ci, ci]
        L.eigen_GPU_batch_bufsize.restype = sz
        L.eigen_GPU_batch_version.restype = ctypes.c_char_p
        ll = ctypes.c_longlong
        ull = ctypes.c_ulonglong
        L.eigen_GPU_warmnet_params.argtypes = [ci, ci]
        L.eigen_GPU_warmnet_params.restype = sz
        L.eigen_GPU_warmnet_bufsize.argtypes = [ci, ci, ci]
        L.eigen_GPU_warmnet_bufsize.restype = sz
        L.eigen_GPU_warmnet_init.argtypes = [ci, ci, vp, ull, vp]
        L.eigen_GPU_warmnet_init.restype = ci
        L.eigen_GPU_warmnet_forward.argtypes = [ci, ci, ci, ci, vp, vp, vp, ci, vp, vp]
        L.eigen_GPU_warmnet_forward.restype = ci
        L.eigen_GPU_warmnet_grad.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp]
        L.eigen_GPU_warmnet_grad.restype = ci
        L.eigen_GPU_warmnet_adam.argtypes = [ll, vp, vp, vp, vp, ci, cf, vp]
        L.eigen_GPU_warmnet_adam.restype = ci
        L.eigen_GPU_batch_refine_bufsize.argtypes = [ci, ci]
        L.eigen_GPU_batch_refine_bufsize.restype = sz
        L.eigen_GPU_batch_refine_DP.argtypes = [ci, ci, ci, vp, vp, ci, vp, vp, vp, ci, vp]
        L.eigen_GPU_batch_refine_DP.restype = ci
        L.eigen_GPU_batch_warm_bufsize.argtypes = [ci, ci, ci]
        L.eigen_GPU_batch_warm_bufsize.restype = sz
        for name in ("eigen_GPU_batch_warm_FP", "eigen_GPU_batch_warm_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, ci, vp, vp]
            f.restype = ci
        for name in ("eigen_GPU_batch_qr_FP", "eigen_GPU_batch_qr_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, vp, vp, vp]
            f.restype = ci
        L.eigen_GPU_batch_interleaved_bufsize.argtypes = [ci, ci, ci, ci, ci, ci]
        L.eigen_GPU_batch_interleaved_bufsize.restype = sz
        for name in ("eigen_GPU_batch_interleaved_FP", "eigen_GPU_batch_interleaved_DP"):
            f = getattr(L, name)
            f.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, ci, vp]
            f.restype = ci
        L.eigen_GPU_batch_gemm.argtypes = [ci, ci, ci, ci, ci, ci, cd, vp, ll, ll, vp, ll, ll, cd, vp, ll, ll,
                                           vp, ll, ll, ci, vp]
        L.eigen_GPU_batch_gemm.restype = ci
        _lib = L
    return _lib


def bufsize(L: int, nm: int, n: int, m: int, dtype_bytes: int, solver: int = IMTQL2) -> int:
    return int(lib().eigen_GPU_batch_bufsize(L, nm, n, m, dtype_bytes, solver))


def query_c_abi(L: int, nm: int, n: int, m: int, dtype_bytes: int) -> int:
    """The reference's C query protocol: wk=NULL, the size_t lands in *w (host memory)."""
    out = ctypes.c_size_t(12345)
    f = lib().eigen_GPU_batch_FP if dtype_bytes == 4 else lib().eigen_GPU_batch_DP
    f(L, nm, n, m, None, ctypes.addressof(out), None, None)
    return int(out.value)


def _stream_ptr(stream) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def eigen_batch(a, w=None, m: int | None = None, solver: int = IMTQL2, info=None,
                workspace=None, stream=None, abi: bool = False, path: str = "auto", phases=None,
                unsorted: bool = False):
    """Solve in place: ``a`` (L, n, nm) float32/float64 CUDA tensor.

    Returns (a, w, info): eigenvectors overwrite ``a`` (column j of matrix k is
    ``a[k, j, :n]``), ``w`` (L, n) ascending eigenvalues, ``info`` (L,) int32
    per-matrix status (0 ok, l>0 QL did not converge) or None when abi=True.
    abi=True calls the reference-compatible eigen_GPU_batch_FP/_DP symbol.
    path selects the n <= 64 implementation: "auto", "fused" (one kernel),
    "split" (lane-per-matrix QL recurrence) or "split-fallback" (split with the
    rotation log shrunk so that every matrix takes the overflow fallback) or
    "split-lds" (split with the QL recurrence's d, e kept in LDS rather than in
    registers) or, for 64 < n <= 256, "split-full" (split with the
    tridiagonalisation on the full matrix instead of its lower triangle) or
    "split-ws" (split with the workspace-resident lower-triangle
    tridiagonalisation instead of the register-resident one).
    phases (split path only): iterable of "tri", "ql", "apply" to run.
    m < n: all n eigenvalues, eigenvectors of the first m only (columns m..n-1
    of each matrix are left unspecified).  unsorted: the QL / Jacobi order.
    """
    import torch
    if not a.is_cuda:
        raise ValueError("eigen_batch needs a device tensor (no CPU path)")
    if a.dtype not in (torch.float32, torch.float64) or a.dim() != 3 or not a.is_contiguous():
        raise ValueError("a must be a contiguous (L, n, nm) float32/float64 tensor")
    L, n, nm = a.shape
    m = n if m is None else m
    tb = a.element_size()
    solver = solver | PATHS[path] | (FLAG_UNSORTED if unsorted else 0)
    if phases:
        for ph in phases:
            solver |= PHASES[ph]
    if w is None:
        w = torch.empty((L, n), dtype=a.dtype, device=a.device)
    if workspace is None:
        workspace = torch.empty(max(bufsize(L, nm, n, m, tb, solver), 1), dtype=torch.uint8,
                                device=a.device)
    sp = _stream_ptr(stream)
    if abi:
        f = lib().eigen_GPU_batch_FP if tb == 4 else lib().eigen_GPU_batch_DP
        f(L, nm, n, m, a.data_ptr(), w.data_ptr(), workspace.data_ptr(), sp)
        return a, w, None
    if info is None:
        info = torch.empty(L, dtype=torch.int32, device=a.device)
    f = lib().eigen_GPU_batch_FP_x if tb == 4 else lib().eigen_GPU_batch_DP_x
    rc = f(L, nm, n, m, a.data_ptr(), w.data_ptr(), workspace.data_ptr(), sp, solver,
           info.data_ptr())
    if rc != 0:
        raise RuntimeError(f"eigen_GPU_batch_{'FP' if tb == 4 else 'DP'}_x failed: hipError {rc}")
    return a, w, info


def refine(a, x, iters: int = 3, w=None, info=None, workspace=None, stream=None):
    """FP64 refinement of approximate eigenvectors (eigen_GPU_batch_refine_DP):
    a (L, n, nm) float64 matrices (upper triangle read), x (L, n, ldx) float64
    approximate eigenvectors (column j of matrix k = x[k, j, :n]), refined in
    place.  Returns (x, w, info): info 0 accepted, else reason bits."""
    import torch
    L, n, nm = a.shape
    ldx = x.shape[2]
    if a.dtype != torch.float64 or x.dtype != torch.float64 or x.shape[:2] != (L, n):
        raise ValueError("float64 tensors a (L, n, nm) and x (L, n, ldx)")
    if w is None:
        w = torch.empty((L, n), dtype=torch.float64, device=a.device)
    if info is None:
        info = torch.empty(L, dtype=torch.int32, device=a.device)
    if workspace is None:
        workspace = torch.empty(max(int(lib().eigen_GPU_batch_refine_bufsize(L, n)), 1), dtype=torch.uint8,
                                device=a.device)
    _check_rc(lib().eigen_GPU_batch_refine_DP(L, nm, n, a.data_ptr(), x.data_ptr(), ldx, w.data_ptr(),
                                              workspace.data_ptr(), _stream_ptr(stream), iters, info.data_ptr()),
              "eigen_GPU_batch_refine_DP")
    return x, w, info


def qr(a, q=None, stream=None):
    """Q factors of the Householder QR of a (L, n, nm) batch (eigen_GPU_batch_qr_*),
    R with a non-negative diagonal; n <= 64.  Returns q (same layout)."""
    import torch
    L, n, nm = a.shape
    if q is None:
        q = torch.empty_like(a)
    f = getattr(lib(), "eigen_GPU_batch_qr_" + _sfx(a))
    _check_rc(f(L, nm, n, a.data_ptr(), q.data_ptr(), _stream_ptr(stream)), "eigen_GPU_batch_qr")
    return q


def eigen_batch_warm(a, q0, max_sweeps: int = 6, m: int | None = None, w=None, info=None, sweeps=None,
                     workspace=None, stream=None):
    """Warm-started solve (eigen_GPU_batch_warm_*): a (L, n, nm) solved in place
    from the approximate eigenbases q0 (same layout), with the plain solver as
    fallback past max_sweeps Jacobi sweeps.  Returns (a, w, info, sweeps)."""
    import torch
    L, n, nm = a.shape
    m = n if m is None else m
    if q0.shape != a.shape or q0.dtype != a.dtype:
        raise ValueError("q0 must have a's shape and dtype")
    if w is None:
        w = torch.empty((L, n), dtype=a.dtype, device=a.device)
    if info is None:
        info = torch.empty(L, dtype=torch.int32, device=a.device)
    if sweeps is None:
        sweeps = torch.empty(L, dtype=torch.int32, device=a.device)
    if workspace is None:
        workspace = torch.empty(max(int(lib().eigen_GPU_batch_warm_bufsize(L, n, a.element_size())), 1),
                                dtype=torch.uint8, device=a.device)
    f = getattr(lib(), "eigen_GPU_batch_warm_" + _sfx(a))
    _check_rc(f(L, nm, n, m, a.data_ptr(), q0.data_ptr(), w.data_ptr(), workspace.data_ptr(), _stream_ptr(stream),
                max_sweeps, info.data_ptr(), sweeps.data_ptr()), "eigen_GPU_batch_warm")
    return a, w, info, sweeps


def eigen_batch_interleaved(a, L: int, nm: int, n: int, m: int | None = None, w=None, solver: int = IMTQL2,
                            info=None, workspace=None, stream=None):
    """Interleaved layout (eigen_GPU_batch_interleaved_*): ``a`` a contiguous
    tensor of L*nm*n elements with element (r, c) of matrix k at
    a[k + L*(r + c*nm)] (e.g. shape (n, nm, L)); eigenvalue j of matrix k goes
    to w[k + L*j].  Solved in place; returns (a, w, info)."""
    import torch
    m = n if m is None else m
    if not a.is_cuda or a.dtype not in (torch.float32, torch.float64) or not a.is_contiguous():
        raise ValueError("a must be a contiguous float32/float64 device tensor")
    if a.numel() != L * nm * n:
        raise ValueError("a must hold L*nm*n elements")
    tb = a.element_size()
    if w is None:
        w = torch.empty((n, L), dtype=a.dtype, device=a.device)
    if info is None:
        info = torch.empty(L, dtype=torch.int32, device=a.device)
    if workspace is None:
        workspace = torch.empty(max(int(lib().eigen_GPU_batch_interleaved_bufsize(L, nm, n, m, tb, solver)), 1),
                                dtype=torch.uint8, device=a.device)
    f = lib().eigen_GPU_batch_interleaved_FP if tb == 4 else lib().eigen_GPU_batch_interleaved_DP
    _check_rc(f(L, nm, n, m, a.data_ptr(), w.data_ptr(), workspace.data_ptr(), _stream_ptr(stream), solver,
                info.data_ptr()), "eigen_GPU_batch_interleaved")
    return a, w, info


UNEVEN_BAD_ORDER = -3  # info value of a matrix whose order is outside 1..nmax


def uneven_bufsize(L: int, nmax: int, dtype_bytes: int, solver: int = IMTQL2) -> int:
    return int(lib().eigen_GPU_batch_uneven_bufsize(L, nmax, dtype_bytes, solver))


def pack_uneven(mats, device="cuda"):
    """Pack square matrices (a list of (n_k, n_k) arrays/tensors, row r column c
    of matrix k at mats[k][r, c]) into the packed uneven layout: returns
    (a, nv) with a 1-D tensor holding matrix k column-major after matrix k-1
    and nv the int32 orders."""
    import torch
    ns = [int(m.shape[0]) for m in mats]
    parts = [torch.as_tensor(m).t().contiguous().reshape(-1) for m in mats]
    a = torch.cat(parts).to(device)
    nv = torch.tensor(ns, dtype=torch.int32, device=device)
    return a, nv


def unpack_uneven(a, w, ns):
    """Inverse of pack_uneven for results: list of (Z_k, w_k) on the host, with
    Z_k[:, j] the eigenvector of w_k[j]."""
    a = a.cpu()
    w = w.cpu()
    out, oa, ow = [], 0, 0
    for n in ns:
        z = a[oa:oa + n * n].reshape(n, n).t().contiguous()
        out.append((z.numpy(), w[ow:ow + n].numpy()))
        oa += n * n
        ow += n
    return out


def eigen_batch_uneven(a, nv, nmax: int, w=None, aoff=None, woff=None, solver: int = IMTQL2,
                       info=None, workspace=None, stream=None):
    """Solve L matrices of different orders in place (the uneven entry points).

    ``a`` 1-D float32/float64 device tensor; ``nv`` (L,) int32 device tensor of
    orders; ``nmax`` a host-side bound on the orders (<= 512).  Without
    ``aoff``/``woff`` (int64 device tensors of element offsets) the layout is
    packed and ``w`` defaults to a tensor of sum(n) elements (one host read of
    nv).  Returns (a, w, info)."""
    import torch
    if not (a.is_cuda and nv.is_cuda):
        raise ValueError("eigen_batch_uneven needs device tensors (no CPU path)")
    if a.dtype not in (torch.float32, torch.float64) or a.dim() != 1 or not a.is_contiguous():
        raise ValueError("a must be a contiguous 1-D float32/float64 tensor")
    if nv.dtype != torch.int32 or nv.dim() != 1 or not nv.is_contiguous():
        raise ValueError("nv must be a contiguous 1-D int32 tensor")
    if (aoff is None) != (woff is None):
        raise ValueError("give both aoff and woff, or neither (packed layout)")
    for o in (aoff, woff):
        if o is not None and (o.dtype != torch.int64 or not o.is_cuda or o.numel() != nv.numel()):
            raise ValueError("aoff/woff must be (L,) int64 device tensors")
    L = nv.numel()
    tb = a.element_size()
    if w is None:
        if aoff is not None:
            raise ValueError("w is required with explicit offsets")
        w = torch.empty(int(nv.clamp(min=0).sum().item()), dtype=a.dtype, device=a.device)
    if info is None:
        info = torch.empty(L, dtype=torch.int32, device=a.device)
    if workspace is None:
        workspace = torch.empty(max(uneven_bufsize(L, nmax, tb, solver), 1), dtype=torch.uint8,
                                device=a.device)
    f = lib().eigen_GPU_batch_uneven_FP if tb == 4 else lib().eigen_GPU_batch_uneven_DP
    rc = f(L, nmax, nv.data_ptr(), None if aoff is None else aoff.data_ptr(),
           None if woff is None else woff.data_ptr(), a.data_ptr(), w.data_ptr(),
           workspace.data_ptr(), _stream_ptr(stream), solver, info.data_ptr())
    if rc != 0:
        raise RuntimeError(f"eigen_GPU_batch_uneven_{'FP' if tb == 4 else 'DP'} failed: hipError {rc}")
    return a, w, info


def check(a_full, w, z, m: int | None = None, stream=None):
    """GPU accuracy check (eigen_GPU_check_x_*): a_full (L, n, nm) full
    symmetric inputs, w (L, n), z (L, n, nm) eigenvectors.  Returns a (L, 4)
    float64 device tensor: err1, err2, TOL, fail bits (1 residual, 2 orth)."""
    import torch
    L, n, nm = a_full.shape
    m = n if m is None else m
    out = torch.empty((L, 4), dtype=torch.float64, device=a_full.device)
    f = lib().eigen_GPU_check_x_FP if a_full.element_size() == 4 else lib().eigen_GPU_check_x_DP
    rc = f(L, nm, n, m, a_full.data_ptr(), w.data_ptr(), z.data_ptr(), out.data_ptr(), _stream_ptr(stream))
    if rc != 0:
        raise RuntimeError(f"eigen_GPU_check_x failed: hipError {rc}")
    return out


def autotune(a_src, solver: int = IMTQL2, stream=None):
    """Run the library's autotuner on the batch a_src (L, n, nm), which is left
    untouched.  Returns (best_bits, best_ms, {bits: ms}, a, w): the chosen
    tuning bits (remembered by the library for later calls of this shape), the
    candidate times, and the eigenpairs of the chosen configuration."""
    import torch
    if not a_src.is_cuda:
        raise ValueError("autotune needs a device tensor")
    L, n, nm = a_src.shape
    tb = a_src.element_size()
    a = torch.empty_like(a_src)
    w = torch.empty((L, n), dtype=a_src.dtype, device=a_src.device)
    ws = torch.empty(max(int(lib().eigen_GPU_batch_tune_bufsize(L, nm, n, tb, solver)), 1),
                     dtype=torch.uint8, device=a_src.device)
    best = ctypes.c_int(0)
    best_ms = ctypes.c_float(0.0)
    cand = (ctypes.c_int * 32)()
    cms = (ctypes.c_float * 32)()
    nc = ctypes.c_int(0)
    f = lib().eigen_GPU_batch_tune_FP if tb == 4 else lib().eigen_GPU_batch_tune_DP
    rc = f(L, nm, n, a_src.data_ptr(), a.data_ptr(), w.data_ptr(), ws.data_ptr(), _stream_ptr(stream),
           solver, ctypes.addressof(best), ctypes.addressof(best_ms), 32, ctypes.addressof(cand),
           ctypes.addressof(cms), ctypes.addressof(nc))
    if rc != 0:
        raise RuntimeError(f"eigen_GPU_batch_tune failed: hipError {rc}")
    return best.value, best_ms.value, {cand[i]: cms[i] for i in range(nc.value)}, a, w


def _sfx(t) -> str:
    import torch
    if t.dtype == torch.float32:
        return "FP"
    if t.dtype == torch.float64:
        return "DP"
    raise ValueError("float32 or float64 tensors only")


def _check_rc(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


def symdmt(z, d, c=None, stream=None):
    """C_k = Z_k diag(d_k) Z_k^T (eigen_GPU_batch_symdmt_*).  z (L, n, nm) in
    the solver's layout (column j of matrix k is z[k, j, :n]), d (L, n).
    Returns c (L, n, nm), exactly symmetric; rows n..nm-1 untouched."""
    import torch
    L, n, nm = z.shape
    if c is None:
        c = torch.empty_like(z)
    if not (z.is_cuda and d.is_cuda and c.is_cuda) or d.dtype != z.dtype or c.dtype != z.dtype:
        raise ValueError("device tensors of one float dtype")
    f = getattr(lib(), "eigen_GPU_batch_symdmt_" + _sfx(z))
    _check_rc(f(L, nm, n, z.data_ptr(), d.data_ptr(), c.data_ptr(), _stream_ptr(stream)), "symdmt")
    return c


def matpow(z, w, p: float, scale: float = 1.0, c=None, stream=None):
    """C_k = scale * Z_k diag(w_k)^p Z_k^T (eigen_GPU_batch_matpow_*), from the
    solver's eigenvectors z (L, n, nm) and eigenvalues w (L, n)."""
    import torch
    L, n, nm = z.shape
    if c is None:
        c = torch.empty_like(z)
    if not (z.is_cuda and w.is_cuda and c.is_cuda) or w.dtype != z.dtype or c.dtype != z.dtype:
        raise ValueError("device tensors of one float dtype")
    f = getattr(lib(), "eigen_GPU_batch_matpow_" + _sfx(z))
    _check_rc(f(L, nm, n, z.data_ptr(), w.data_ptr(), p, scale, c.data_ptr(), _stream_ptr(stream)), "matpow")
    return c


def gram(y, p: int, alpha: float = 0.0, rinv=None, rstride: int = 0, nm: int | None = None, a=None, stream=None):
    """A_k = alpha I + Y_k^T diag(r_k) Y_k (eigen_GPU_batch_gram_*).  y (L, n,
    ldy): y[k, i, o] = Y_k(o, i), o < p <= ldy.  rinv: None (R = I) or a device
    tensor with r_k = rinv.flatten()[k * rstride:][:p].  Returns a (L, n, nm)."""
    import torch
    L, n, ldy = y.shape
    nm = n if nm is None else nm
    if a is None:
        a = torch.empty((L, n, nm), dtype=y.dtype, device=y.device)
    if rinv is not None and rinv.dtype != y.dtype:
        raise ValueError("rinv must have y's dtype")
    f = getattr(lib(), "eigen_GPU_batch_gram_" + _sfx(y))
    _check_rc(f(L, nm, n, p, y.data_ptr(), ldy, None if rinv is None else rinv.data_ptr(), rstride, alpha,
                a.data_ptr(), _stream_ptr(stream)), "gram")
    return a


def tune_set(L: int, n: int, dtype_bytes: int, bits: int, solver: int = IMTQL2) -> None:
    """Remember tuning bits for a shape (bits = 0 forgets it); see eigen_GPU_batch_tune_set."""
    _check_rc(lib().eigen_GPU_batch_tune_set(L, n, dtype_bytes, solver, bits), "eigen_GPU_batch_tune_set")


def launch_params(L: int, n: int, dtype_bytes: int, solver: int = IMTQL2) -> dict:
    """Launch shape of the first kernel a call takes (eigen_GPU_batch_launch_params)."""
    p = (ctypes.c_int * 6)()
    rc = lib().eigen_GPU_batch_launch_params(L, n, dtype_bytes, solver, ctypes.addressof(p))
    keys = ("grpSZ", "numTB", "numGR", "numTH", "sizeSH", "sizeL2")
    out = dict(zip(keys, list(p)))
    out["rc"] = rc
    return out


def gemm_raw(kind: int, transa: int, transb: int, M: int, N: int, K: int, alpha: float,
             A, lda: int, sa: int, B, ldb: int, sb: int, beta: float, D, ldd: int, sd: int,
             C, ldc: int, sc: int, batch: int, stream=None) -> int:
    """eigen_GPU_batch_gemm on device tensors (their data pointers): C_b = alpha
    op(A_b) op(B_b) + beta D_b, column-major, strides in elements.  Returns the
    hipError code."""
    return int(lib().eigen_GPU_batch_gemm(kind, transa, transb, M, N, K, alpha, A.data_ptr(), lda, sa,
                                          B.data_ptr(), ldb, sb, beta,
                                          None if D is None else D.data_ptr(), ldd, sd,
                                          C.data_ptr(), ldc, sc, batch, _stream_ptr(stream)))


def gemm(a, b, transa: bool = False, transb: bool = False, kind: int | None = None, alpha: float = 1.0,
         beta: float = 0.0, d=None, c=None, stream=None):
    """Batched product in the solver's layout: a (batch, cols, ld) holds the
    column-major matrices A_b (element (r, c) at a[b, c, r]); likewise b, d, c.
    Returns c with C_b = alpha op(A_b) op(B_b) + beta D_b (shape (batch, N, M)).
    kind defaults to GEMM_F64 for float64 and GEMM_F32 for float32."""
    import torch
    if a.dtype != b.dtype or a.dim() != 3 or b.dim() != 3 or not (a.is_contiguous() and b.is_contiguous()):
        raise ValueError("contiguous (batch, cols, ld) tensors of one dtype")
    if kind is None:
        kind = GEMM_F64 if a.dtype == torch.float64 else GEMM_F32
    batch = a.shape[0]
    ar, ac = a.shape[2], a.shape[1]
    br, bc = b.shape[2], b.shape[1]
    M, K = (ac, ar) if transa else (ar, ac)
    K2, N = (bc, br) if transb else (br, bc)
    if K != K2 or b.shape[0] != batch:
        raise ValueError("inner dimensions or batch sizes differ")
    if c is None:
        c = torch.empty((batch, N, M), dtype=a.dtype, device=a.device)
    rc = gemm_raw(kind, int(transa), int(transb), M, N, K, alpha, a, ar, ar * ac, b, br, br * bc, beta,
                  d, M if d is None else d.shape[2], 0 if d is None else d.shape[1] * d.shape[2],
                  c, c.shape[2], c.shape[1] * c.shape[2], batch, stream)
    _check_rc(rc, "eigen_GPU_batch_gemm")
    return c


def tune_reset() -> None:
    lib().eigen_GPU_batch_tune_reset()


def version() -> str:
    return lib().eigen_GPU_batch_version().decode()
