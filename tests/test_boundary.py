"""The drop-in boundary, checked without a GPU.

* libeigenGbatch.so loads and exports every function include/*.h declares;
* the reference's workspace-query protocol (wk == NULL writes a size_t into the
  host memory at w, eigen_GPU_batch.cu:329-359) and its validation rule
  (L<1, nm<n, n<1, m<1, m>n -> no-op, eigen_GPU_batch.cu:252);
* the C++ templates of include/eigen_GPU_batch.hpp compile against the library
  and report the same workspace size (compiled and run on the host only: the
  query makes no HIP call).
"""
import ctypes
import os
import re
import subprocess

import pytest

import eigeng_batched as eg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "eigeng-batched_amd")


def header_functions():
    names = set()
    for f in os.listdir(INC):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(INC, f)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(eigen_GPU_batch\w*)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert {"eigen_GPU_batch_FP", "eigen_GPU_batch_DP"} <= set(names)
    lib = ctypes.CDLL(eg.LIB_PATH)
    for nme in names:
        assert hasattr(lib, nme), f"{nme} declared in include/ but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", eg.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for nme in names:
        assert re.search(rf"\bT {nme}$", out, flags=re.M), nme


def test_static_archive_has_the_abi():
    a = os.path.join(PKG, "libeigenGbatch.a")
    assert os.path.exists(a)
    out = subprocess.run(["nm", "--defined-only", a], capture_output=True, text=True, check=True).stdout
    for nme in ("eigen_GPU_batch_FP", "eigen_GPU_batch_DP"):
        assert re.search(rf"\bT {nme}$", out, flags=re.M)


@pytest.mark.parametrize("tb", [4, 8])
def test_query_protocol(tb):
    for n in (1, 2, 17, 32, 64):
        lw = eg.query_c_abi(100, n + (n & 1), n, n, tb)
        assert lw > 0, "query must give a nonzero size so that wk != NULL after allocation"
        assert lw == eg.bufsize(100, n + (n & 1), n, n, tb)
    # invalid arguments: size 0, as eigen_GPU_batch_BufferSize returns with lwork = 0
    for (L, nm, n, m) in [(0, 4, 4, 4), (4, 3, 4, 4), (4, 4, 0, 1), (4, 4, 4, 0), (4, 4, 4, 5)]:
        assert eg.query_c_abi(L, nm, n, m, tb) == 0
        assert eg.bufsize(L, nm, n, m, tb) == 0


def test_query_writes_a_full_size_t():
    buf = (ctypes.c_ubyte * 16)(*([0xAB] * 16))
    eg.lib().eigen_GPU_batch_FP(10, 8, 8, 8, None, ctypes.addressof(buf), None, None)
    val = ctypes.c_size_t.from_buffer(buf).value
    assert val == eg.bufsize(10, 8, 8, 8, 4)
    assert bytes(buf[8:]) == bytes([0xAB] * 8), "query must not write past the size_t"


def test_version():
    assert eg.version().startswith("eigen-mi355x")


def test_cpp_template_api(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include "eigen_GPU_batch.hpp"
#include <cstdio>
int main() {
  size_t lw_f = 0, lw_d = 0, lw_bad = 99;
  gpuError_t e1 = eigen_GPU_batch_BufferSize<float>(16, 64, 64, 64, (float*)0, (float*)0, &lw_f);
  gpuError_t e2 = eigen_GPU_batch_BufferSize<double>(16, 33, 33, 33, (double*)0, (double*)0, &lw_d);
  gpuError_t e3 = eigen_GPU_batch_BufferSize<float>(16, 8, 9, 9, (float*)0, (float*)0, &lw_bad);
  size_t q = 0;
  eigen_GPU_batch_FP(16, 64, 64, 64, 0, (float*)&q, 0, 0);
  std::printf("%d %d %d %zu %zu %zu %zu\n", (int)e1, (int)e2, (int)e3, lw_f, lw_d, lw_bad, q);
  return 0;
}
''')
    exe = tmp_path / "t"
    # plain host C++ (g++), as a caller's host code would be
    cmd = ["g++", "-std=c++14", "-D__HIP_PLATFORM_AMD__", f"-I{INC}", "-I/opt/rocm/include", str(src),
           "-o", str(exe), f"-L{PKG}", "-leigenGbatch", f"-Wl,-rpath,{PKG}"]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    e1, e2, e3, lw_f, lw_d, lw_bad, q = map(int, out)
    assert e1 == 0 and e2 == 0 and e3 != 0
    assert lw_f == eg.bufsize(16, 64, 64, 64, 4) and lw_d == eg.bufsize(16, 33, 33, 33, 8)
    assert lw_bad == 0 and q == lw_f


@pytest.mark.parametrize("sfx", ["FP", "DP"])
def test_products_validate_arguments_without_a_gpu(sfx):
    """The product entry points reject bad sizes and null pointers before any
    HIP call (hipErrorInvalidValue = 1), and L == 0 is a successful no-op."""
    lib = eg.lib()
    sym = getattr(lib, "eigen_GPU_batch_symdmt_" + sfx)
    mp = getattr(lib, "eigen_GPU_batch_matpow_" + sfx)
    gr = getattr(lib, "eigen_GPU_batch_gram_" + sfx)
    z, d, c = 4096, 8192, 1 << 20                  # never dereferenced: every call returns first
    for L, nm, n in ((1, 4, 0), (1, 3, 4), (-1, 4, 4), (1, 513, 513)):
        assert sym(L, nm, n, z, d, c, None) == 1
        assert mp(L, nm, n, z, d, -0.5, 1.0, c, None) == 1
        assert gr(L, nm, n, 8, z, 8, None, 0, 1.0, c, None) == 1
    assert sym(1, 4, 4, z, d, None, None) == 1      # null output
    assert gr(1, 4, 4, 8, z, 7, None, 0, 1.0, c, None) == 1     # ldy < p
    assert gr(1, 4, 4, 8, z, 8, None, -1, 1.0, c, None) == 1    # negative rinv stride
    assert sym(0, 4, 4, z, d, c, None) == 0
    assert gr(0, 4, 4, 8, z, 8, None, 0, 1.0, c, None) == 0


def test_uneven_workspace_sizes():
    """The uneven workspace: zero for invalid arguments; beyond order 64 it
    holds the split route's slots (chunks of 65536 FP32 / 8192 FP64 matrices)."""
    assert eg.uneven_bufsize(0, 64, 4) == 0 and eg.uneven_bufsize(10, 0, 4) == 0
    assert eg.uneven_bufsize(10, 513, 4) == 0 and eg.uneven_bufsize(10, 64, 2) == 0
    small = eg.uneven_bufsize(100000, 64, 4)
    big = eg.uneven_bufsize(100000, 128, 4)
    assert small < 64 << 20 < big
    # slots of order 128: W, rotation log, sweep log, d/e, betas, output slots
    per = 128 * 128 * 4 + (3 * 128 * 128 // 2 + 16) * 8 + (3 * 128 + 8) * 4 + 2 * 128 * 4 + 128 * 4 + 128 * 4 + 8
    assert 65536 * per <= big <= 65536 * per + small + (256 * 128 * 128 * 4) + (64 << 10)
    assert eg.uneven_bufsize(1000, 128, 4) < big                 # chunks never exceed the batch
    assert eg.uneven_bufsize(100000, 200, 8) > eg.uneven_bufsize(100000, 64, 8)
