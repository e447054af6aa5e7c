"""GPU: the batched matrix-core products (eigen_GPU_batch_gemm) against float64
torch products of the same (rounded) operands.

Bars: FP64 operands, |C - C_ref| <= 4 K eps64 (|A| |B|)_ij + tiny; FP32 operands
(v_mfma_f32_16x16x4f32) likewise with eps32; FP16 / BF16 operands: the reference
is formed from the operands rounded to half / bfloat16 (what the matrix cores
multiply), accumulation in FP32, so the same bound with eps32.
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bf16_round(x):
    return x.to(torch.bfloat16).to(x.dtype)


def ref_and_bound(A, B, ta, tb, kind):
    """A, B: (batch, cols, ld) column-major tensors. Returns (C_ref, bound) as
    float64 (batch, N, M) tensors in the same layout."""
    a = A.double()
    b = B.double()
    if kind == eg.GEMM_F16:
        a, b = A.half().double(), B.half().double()
    elif kind == eg.GEMM_BF16:
        a, b = bf16_round(A).double(), bf16_round(B).double()
    elif kind == eg.GEMM_F16X3:
        # the split operands: hi + lo / 2^11, both halves rounded
        def split(X):
            hi = X.half().float()
            lo = ((X - hi) * 2048.0).half().float()
            return hi.double() + lo.double() / 2048.0
        a, b = split(A), split(B)
    # row-major views of the column-major matrices: Am[b] = a[b].T
    Am = a.transpose(1, 2)
    Bm = b.transpose(1, 2)
    opA = Am.transpose(1, 2) if ta else Am
    opB = Bm.transpose(1, 2) if tb else Bm
    C = torch.bmm(opA.contiguous(), opB.contiguous())
    Cabs = torch.bmm(opA.abs().contiguous(), opB.abs().contiguous())
    K = opA.shape[2]
    eps = 2.0 ** -52 if kind == eg.GEMM_F64 else 2.0 ** -23
    bound = 4.0 * K * eps * Cabs + 1e-300
    if kind == eg.GEMM_F16X3:
        # the dropped lo*lo term: 2^-22 of |A||B| per product, plus FP32 accumulation
        bound = bound + 2.0 ** -21 * Cabs
    return C.transpose(1, 2), bound.transpose(1, 2)


KINDS = [eg.GEMM_F64, eg.GEMM_F32, eg.GEMM_F16, eg.GEMM_BF16, eg.GEMM_F16X3]
SHAPES = [(1, 1, 1), (17, 33, 5), (64, 64, 64), (100, 70, 130), (256, 256, 256), (5, 200, 48)]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape", SHAPES)
def test_gemm_vs_torch(kind, ta, tb, shape):
    M, N, K = shape
    print(f"[gemm] kind={kind} ta={ta} tb={tb} M={M} N={N} K={K}", flush=True)
    dt = torch.float64 if kind == eg.GEMM_F64 else torch.float32
    g = torch.Generator(device="cpu").manual_seed(M * 131 + N * 7 + K + 1000 * kind + 10 * ta + tb)
    batch = 3
    # column-major storage: A is (M x K) or (K x M) when transposed
    ar, ac = (K, M) if ta else (M, K)
    br, bc = (N, K) if tb else (K, N)
    A = torch.randn((batch, ac, ar), generator=g, dtype=torch.float64).to(dt).to(DEV)
    B = torch.randn((batch, bc, br), generator=g, dtype=torch.float64).to(dt).to(DEV)
    C = eg.gemm(A, B, transa=bool(ta), transb=bool(tb), kind=kind)
    torch.cuda.synchronize()
    Cr, bound = ref_and_bound(A, B, ta, tb, kind)
    err = (C.double() - Cr).abs()
    assert C.shape == (batch, N, M)
    assert bool((err <= bound).all()), f"max err/bound {float((err / bound).max()):.3f}"


@pytest.mark.parametrize("kind", [eg.GEMM_F64, eg.GEMM_F32])
def test_gemm_alpha_beta_in_place(kind):
    """C = alpha A B + beta C with D == C, and with a separate D."""
    dt = torch.float64 if kind == eg.GEMM_F64 else torch.float32
    g = torch.Generator(device="cpu").manual_seed(5)
    batch, n = 4, 96
    A = torch.randn((batch, n, n), generator=g, dtype=torch.float64).to(dt).to(DEV)
    B = torch.randn((batch, n, n), generator=g, dtype=torch.float64).to(dt).to(DEV)
    C0 = torch.randn((batch, n, n), generator=g, dtype=torch.float64).to(dt).to(DEV)
    C = C0.clone()
    eg.gemm(A, B, kind=kind, alpha=-0.5, beta=2.0, d=C, c=C)
    D = C0.clone()
    C2 = torch.empty_like(C0)
    eg.gemm(A, B, kind=kind, alpha=-0.5, beta=2.0, d=D, c=C2)
    torch.cuda.synchronize()
    Cr, bound = ref_and_bound(A, B, 0, 0, kind)
    ref = -0.5 * Cr + 2.0 * C0.double()
    b = 0.5 * bound + 4 * (2.0 ** -23 if kind == eg.GEMM_F32 else 2.0 ** -52) * 2 * C0.double().abs()
    assert bool(((C.double() - ref).abs() <= b).all())
    assert torch.equal(C, C2)


def test_gemm_invalid_arguments():
    x = torch.zeros((2, 8, 8), dtype=torch.float64, device=DEV)
    y = torch.zeros((2, 8, 8), dtype=torch.float64, device=DEV)
    # C overlapping A
    assert eg.gemm_raw(eg.GEMM_F64, 0, 0, 8, 8, 8, 1.0, x, 8, 64, y, 8, 64, 0.0, None, 8, 0, x, 8, 64, 2) != 0
    # bad leading dimension
    assert eg.gemm_raw(eg.GEMM_F64, 0, 0, 8, 8, 8, 1.0, x, 4, 64, y, 8, 64, 0.0, None, 8, 0, y, 8, 64, 1) != 0
    # zero batch is a no-op
    assert eg.gemm_raw(eg.GEMM_F64, 0, 0, 8, 8, 8, 1.0, x, 8, 64, y, 8, 64, 0.0, None, 8, 0, y, 8, 64, 0) == 0
