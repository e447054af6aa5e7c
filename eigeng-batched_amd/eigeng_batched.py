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
# tuning flags OR-ed into the solver id of the _x entry points
FLAG_FUSED, FLAG_SPLIT, FLAG_TINYLOG = 0x100, 0x200, 0x400
PATHS = {"auto": 0, "fused": FLAG_FUSED, "split": FLAG_SPLIT, "split-fallback": FLAG_SPLIT | FLAG_TINYLOG}
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
        L.eigen_GPU_batch_bufsize.argtypes = [ci, ci, ci, ci, ci, ci]
        L.eigen_GPU_batch_bufsize.restype = sz
        L.eigen_GPU_batch_version.restype = ctypes.c_char_p
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
                workspace=None, stream=None, abi: bool = False, path: str = "auto"):
    """Solve in place: ``a`` (L, n, nm) float32/float64 CUDA tensor.

    Returns (a, w, info): eigenvectors overwrite ``a`` (column j of matrix k is
    ``a[k, j, :n]``), ``w`` (L, n) ascending eigenvalues, ``info`` (L,) int32
    per-matrix status (0 ok, l>0 QL did not converge) or None when abi=True.
    abi=True calls the reference-compatible eigen_GPU_batch_FP/_DP symbol.
    path selects the n <= 64 implementation: "auto", "fused" (one kernel),
    "split" (lane-per-matrix QL recurrence) or "split-fallback" (split with the
    rotation log shrunk so that every matrix takes the overflow fallback).
    """
    import torch
    if not a.is_cuda:
        raise ValueError("eigen_batch needs a device tensor (no CPU path)")
    if a.dtype not in (torch.float32, torch.float64) or a.dim() != 3 or not a.is_contiguous():
        raise ValueError("a must be a contiguous (L, n, nm) float32/float64 tensor")
    L, n, nm = a.shape
    m = n if m is None else m
    tb = a.element_size()
    solver = solver | PATHS[path]
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


def version() -> str:
    return lib().eigen_GPU_batch_version().decode()
