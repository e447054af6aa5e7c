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
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(eigen_GPU_\w*)\s*\(", txt, flags=re.M):
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
  int g = -1, tb = -1, gr = -1, th = -1, sh = -1, l2 = -1;
  gpuError_t e4 = eigen_GPU_batch_get_Launch_params<float>(16384, 64, g, tb, gr, th, sh, l2);
  int g2 = -1, tb2 = -1, gr2 = -1, th2 = -1, sh2 = -1, l22 = -1;
  gpuError_t e5 = eigen_GPU_batch_get_Launch_params<double>(0, 64, g2, tb2, gr2, th2, sh2, l22);
  std::printf("%d %d %d %zu %zu %zu %zu %d %d %d %d %d %d %d %d\n", (int)e1, (int)e2, (int)e3, lw_f, lw_d,
              lw_bad, q, (int)e4, g, tb, gr, th, sh, (int)e5, g2 + tb2 + gr2 + th2 + sh2 + l22);
  return 0;
}
''')
    exe = tmp_path / "t"
    # plain host C++ (g++), as a caller's host code would be
    cmd = ["g++", "-std=c++14", "-D__HIP_PLATFORM_AMD__", f"-I{INC}", "-I/opt/rocm/include", str(src),
           "-o", str(exe), f"-L{PKG}", "-leigenGbatch", f"-Wl,-rpath,{PKG}"]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    e1, e2, e3, lw_f, lw_d, lw_bad, q, e4, g, tb, gr, th, sh, e5, zero = map(int, out)
    assert e1 == 0 and e2 == 0 and e3 != 0
    assert lw_f == eg.bufsize(16, 64, 64, 64, 4) and lw_d == eg.bufsize(16, 33, 33, 33, 8)
    assert lw_bad == 0 and q == lw_f
    # get_Launch_params<T>: the split path's first kernel at L = 16384, n = 64
    # (one 64-lane group per matrix, one matrix per one-wave workgroup)
    assert e4 == 0 and (g, tb, gr, th, sh) == (64, 16384, 1, 64, 0)
    assert e5 != 0 and zero == 0       # invalid L: every output zeroed, as the reference does


def test_launch_params_follow_the_dispatch():
    """eigen_GPU_batch_launch_params reports the first kernel of the path a
    call takes: the fused register kernel for small batches (groups of 8..64
    lanes, 64 / group-size matrices per wave), the split path from 2048
    matrices (from one matrix for FP64 16 <= n <= 32 and FP32 28 <= n <= 32,
    from 1024 for FP32 n >= 56), the two-wave FP32
    kernels up to n = 128, the workgroup kernels beyond."""
    eg.tune_reset()
    p = eg.launch_params(1024, 32, 8)          # n = 32 FP64 L = 1024: split, 2 matrices per wave
    assert p["rc"] == 0 and (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (32, 512, 2, 64)
    p = eg.launch_params(1024, 48, 8)          # n = 48 FP64 L = 1024: fused, one matrix per wave
    assert p["rc"] == 0 and (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (64, 1024, 1, 64)
    assert eg.bufsize(1024, 32, 32, 32, 8) > eg.bufsize(1024, 48, 48, 48, 8)   # only split needs a log
    p = eg.launch_params(1024, 8, 8)           # n = 8: fused below 2048 matrices
    assert p["rc"] == 0 and (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (8, 128, 8, 64)
    assert eg.bufsize(1024, 8, 8, 8, 8) < eg.bufsize(2048, 8, 8, 8, 8)
    p = eg.launch_params(100, 5, 4)
    assert (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (8, 13, 8, 64)
    p = eg.launch_params(8192, 128, 4)         # FP32 n = 128, L >= 2048: two-wave register split
    assert (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (128, 8192, 1, 128)
    p = eg.launch_params(8192, 256, 8)         # FP64 n = 256: large split, 256-thread tridiagonalisation
    assert (p["grpSZ"], p["numTB"], p["numGR"], p["numTH"]) == (256, 1024, 1, 256)
    p = eg.launch_params(100, 300, 8)          # small batch beyond 256: workgroup kernel
    assert (p["numTB"], p["numGR"], p["numTH"]) == (100, 1, 256)
    for bad in ((0, 10, 4), (10, 0, 4), (10, 513, 8), (10, 10, 2)):
        p = eg.launch_params(*bad)
        assert p["rc"] != 0 and all(p[k] == 0 for k in ("grpSZ", "numTB", "numGR", "numTH"))


@pytest.mark.parametrize("tb", [4, 8])
def test_tuning_never_changes_the_size_query(tb):
    """A remembered tuning choice never needs more workspace than the default
    size query reports (the reference ABI carries no workspace length): the
    query is the same before and after, and a choice that does not fit is not
    taken (the launch shape stays the default one), while one that fits is."""
    eg.tune_reset()
    try:
        L, n = 1000, 48                         # default: fused (256 bytes); split needs tens of MB
        q0 = eg.query_c_abi(L, n, n, n, tb)
        b0 = eg.bufsize(L, n, n, n, tb)
        p0 = eg.launch_params(L, n, tb)
        assert eg.bufsize(L, n, n, n, tb, eg.FLAG_SPLIT) > q0
        eg.tune_set(L, n, tb, eg.FLAG_SPLIT)
        assert eg.query_c_abi(L, n, n, n, tb) == q0 and eg.bufsize(L, n, n, n, tb) == b0
        assert eg.launch_params(L, n, tb) == p0, "a tuned path that needs more workspace was taken"
        # n = 200, L = 4096: default the large split path (256-thread
        # tridiagonalisation); the 512-thread workgroup kernel needs less
        L2, n2, wg512 = 4096, 200, eg.FLAG_FUSED | (2 << 16)
        q2 = eg.query_c_abi(L2, n2, n2, n2, tb)
        assert eg.launch_params(L2, n2, tb)["numTH"] == 256
        assert eg.bufsize(L2, n2, n2, n2, tb, wg512) < q2
        eg.tune_set(L2, n2, tb, wg512)
        assert eg.query_c_abi(L2, n2, n2, n2, tb) == q2
        p2 = eg.launch_params(L2, n2, tb)
        assert (p2["grpSZ"], p2["numTB"], p2["numGR"], p2["numTH"]) == (512, 1024, 1, 512)
        eg.tune_set(L2, n2, tb, 0)              # forget it again
        assert eg.launch_params(L2, n2, tb)["numTH"] == 256
        with pytest.raises(RuntimeError):
            eg.tune_set(L2, n2, tb, eg.FLAG_SPLIT | 0x1000)    # phase bits cannot be remembered
    finally:
        eg.tune_reset()


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


def test_gemm_validates_arguments_without_a_gpu():
    """eigen_GPU_batch_gemm rejects bad shapes, leading dimensions, kinds and
    overlapping outputs before any HIP call (hipErrorInvalidValue = 1);
    batch == 0 is a successful no-op."""
    f = eg.lib().eigen_GPU_batch_gemm
    A, B, C = 1 << 20, 2 << 20, 3 << 20               # never dereferenced
    ok = dict(kind=0, ta=0, tb=0, M=8, N=8, K=8, A=A, lda=8, sa=64, B=B, ldb=8, sb=64, C=C, ldc=8, sc=64, batch=0)

    def call(**kw):
        a = dict(ok, **kw)
        return f(a["kind"], a["ta"], a["tb"], a["M"], a["N"], a["K"], 1.0, a["A"], a["lda"], a["sa"],
                 a["B"], a["ldb"], a["sb"], 0.0, None, 0, 0, a["C"], a["ldc"], a["sc"], a["batch"], None)
    assert call() == 0
    for bad in (dict(kind=5), dict(kind=-1), dict(M=0), dict(K=0), dict(batch=-1), dict(lda=7, batch=1),
                dict(ta=1, lda=7, batch=1), dict(ldb=7, batch=1), dict(ldc=7, batch=1),
                dict(C=A + 8, batch=1), dict(B=C, batch=1), dict(sa=-1, batch=1)):
        assert call(**bad) == 1, bad


def test_mixed_workspace_sizes():
    """The mixed-precision solver: FP64 interface and n <= 256 only; its
    workspace holds the FP32 copy, the FP32 solver's workspace and one
    refinement chunk (at most 2048 matrices) of six n x n FP64 arrays."""
    assert eg.lib().eigen_GPU_batch_refine_bufsize(10, 257) == 0
    assert eg.lib().eigen_GPU_batch_refine_bufsize(0, 16) == 0
    assert eg.lib().eigen_GPU_batch_warm_bufsize(10, 65, 4) == 0
    assert eg.lib().eigen_GPU_batch_warm_bufsize(100, 48, 4) >= 3 * 100 * 48 * 48 * 4
    assert eg.bufsize(100, 64, 64, 64, 4, eg.MIXED) == 0
    assert eg.bufsize(100, 257, 257, 257, 8, eg.MIXED) == 0
    n = 256
    small = eg.bufsize(100, n, n, n, 8, eg.MIXED)
    big = eg.bufsize(8192, n, n, n, 8, eg.MIXED)
    assert small > 100 * 6 * n * n * 8
    fp32 = eg.bufsize(8192, n, n, n, 4, eg.IMTQL2)
    assert big >= fp32 + 8192 * n * n * 4 + 2048 * 6 * n * n * 8
    # plus the eigenvalues / status of the FP32 solve and the refinement's per-matrix results
    assert big <= fp32 + 8192 * n * n * 4 + 2048 * 6 * n * n * 8 + 8192 * (n + 1) * 8 + 2048 * (160 * n + 64) + (1 << 20)


def test_chunk_bits_shrink_the_workspace():
    """Bits 18-19 divide the split paths' chunk: the workspace shrinks by the
    same factor (to the reference's lwork at n = 256 FP64 with a chunk of 2048,
    eigen_GPU_batch.hpp:114-116 asks 4.31 GB there)."""
    b = [eg.bufsize(16384, 64, 64, 64, 4, sh << eg.CHUNK_SHIFT) for sh in range(4)]
    for s in range(1, 4):
        assert abs(b[s] * (1 << s) - b[0]) <= (1 << s) * (1 << 16)
    assert b[2] < 300e6 and b[3] < 150e6
    d = [eg.bufsize(8192, 256, 256, 256, 8, sh << eg.CHUNK_SHIFT) for sh in range(4)]
    assert d[1] < 8.7e9 and d[2] < 4.4e9
    m = [eg.bufsize(8192, 256, 256, 256, 8, eg.MIXED | (sh << eg.CHUNK_SHIFT)) for sh in range(4)]
    assert m[0] > m[1] > m[2] > m[3]
