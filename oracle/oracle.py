"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by bench.py for its
``cpu_baseline`` leg and by ``__graft_entry__.smoke()`` as the checker.  The GPU
library never imports, links or falls back to anything in this directory.

Layout is the reference's: matrix k at ``a[k*nm*n:]`` column-major with leading
dimension ``nm`` (eigen_GPU_batch.cu:74-90), eigenvalues at ``w[k*n:]``.  Arrays
passed here are numpy arrays of shape (L, n, nm) in C order, i.e. ``a[k, c, r]``
is element (r, c) of matrix k.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

IMTQL2, TQL2, JACOBI = 0, 1, 2
# CPU only: the classical (max-pivot) Jacobi of the reference's jac.cpp:60-217
JACOBI_CLASSICAL = 3
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        for sfx, ct in (("_s", ctypes.c_float), ("_d", ctypes.c_double)):
            p = ctypes.POINTER(ct)
            f = getattr(_lib, "oracle_eig_batch" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, p, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int), ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(_lib, "oracle_tred1" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, p, p, p]
            f.restype = None
            f = getattr(_lib, "oracle_ql" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, p, p, p, ctypes.c_int]
            f.restype = ctypes.c_int
            f = getattr(_lib, "oracle_check" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, p, p, p,
                          ctypes.POINTER(ctypes.c_double)]
            f.restype = None
        _lib.oracle_max_threads.restype = ctypes.c_int
    return _lib


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "_s"
    if dt == np.float64:
        return "_d"
    raise TypeError(f"oracle supports float32/float64, got {dt}")


def _ptr(x: np.ndarray):
    ct = ctypes.c_float if x.dtype == np.float32 else ctypes.c_double
    return x.ctypes.data_as(ctypes.POINTER(ct))


def eig_batch(a: np.ndarray, solver: int = IMTQL2, threads: int = 0):
    """Full pipeline on a copy of ``a`` (L, n, nm): returns (vectors, w, info).

    vectors has the same (L, n, nm) layout (column j = eigenvector of w[:, j]),
    w is (L, n) ascending, info (L,) int32 is the QL error per matrix (0 = ok).
    """
    a = np.ascontiguousarray(a).copy()
    L, n, nm = a.shape
    w = np.zeros((L, n), dtype=a.dtype)
    info = np.zeros(L, dtype=np.int32)
    f = getattr(lib(), "oracle_eig_batch" + _sfx(a.dtype))
    rc = f(L, nm, n, _ptr(a), _ptr(w), solver,
           info.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), threads)
    if rc < 0:
        raise ValueError("oracle_eig_batch: invalid arguments")
    return a, w, info


def tred1(a: np.ndarray):
    """Tridiagonal stage on a copy of one matrix (n, nm): returns (storage, d, e)."""
    a = np.ascontiguousarray(a).copy()
    n, nm = a.shape
    d = np.zeros(n, dtype=a.dtype)
    e = np.zeros(n, dtype=a.dtype)
    getattr(lib(), "oracle_tred1" + _sfx(a.dtype))(nm, n, _ptr(a), _ptr(d), _ptr(e))
    return a, d, e


def ql(d: np.ndarray, e: np.ndarray, variant: int = IMTQL2):
    """QL on the tridiagonal (d, e) with e[n-1] = 0: returns (d, Z, ierr) unsorted."""
    d = np.ascontiguousarray(d).copy()
    e = np.ascontiguousarray(e).copy()
    n = d.shape[0]
    z = np.zeros((n, n), dtype=d.dtype)
    rc = getattr(lib(), "oracle_ql" + _sfx(d.dtype))(n, n, _ptr(d), _ptr(e), _ptr(z), variant)
    return d, z, rc


def check(a0: np.ndarray, w: np.ndarray, z: np.ndarray, m: int | None = None) -> np.ndarray:
    """eigen_GPU_check.cu:34-81 metrics for one matrix.

    a0: original full symmetric matrix (n, nm); w: (n,) eigenvalues; z: (n, nm)
    eigenvectors.  Returns float64[6] = err1, err2, TOL, fail flags,
    max per-pair residual / ||A||_F, err1 / ||A||_F.
    """
    a0 = np.ascontiguousarray(a0)
    n, nm = a0.shape
    z = np.ascontiguousarray(z, dtype=a0.dtype)
    w = np.ascontiguousarray(w, dtype=a0.dtype)
    out = np.zeros(6, dtype=np.float64)
    getattr(lib(), "oracle_check" + _sfx(a0.dtype))(
        nm, n, n if m is None else m, _ptr(a0), _ptr(w), _ptr(z),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
