"""GPU accuracy checker (eigen_GPU_check_FP/_DP and the _x extension) against
the oracle's checker (the reference's check_, eigen_GPU_check.cu:34-81)."""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import workloads as wl

pytestmark = pytest.mark.gpu

DEV = "cuda"


def full_sym(a0):
    """(L, n, nm) upper-triangle batch -> full symmetric copy (padding kept)."""
    full = a0.copy()
    L, n, nm = a0.shape
    up = wl.unpack(np.nan_to_num(a0))       # (L, n, n) symmetric, row-major
    full[:, :, :n] = np.transpose(up, (0, 2, 1))
    return full


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 2, 7, 32, 64, 100, 256])
def test_check_matches_oracle(n, dtype):
    print(f"[check] n={n} dtype={np.dtype(dtype).name}", flush=True)
    nm = n + (n & 1)
    L = 9
    a0 = wl.sym_rand(L, n, nm, dtype=dtype, seed=40 + n)
    a_full = np.nan_to_num(full_sym(a0))
    t = torch.from_numpy(np.nan_to_num(a0)).to(DEV)
    z, w, info = eg.eigen_batch(t)
    # plant failures: matrix 1 gets a wrong eigenvalue, matrix 2 a scaled vector
    w[1, 0] += 1e-2 * (1.0 + float(w[1].abs().max()))
    z[2, 0, :n] *= 1.01
    af = torch.from_numpy(a_full).to(DEV)
    out = eg.check(af, w, z).cpu().numpy()
    zh, wh = z.cpu().numpy(), w.cpu().numpy()
    for k in range(L):
        ref = oracle.check(a_full[k].copy(), wh[k], zh[k])
        # both in double; the GPU contracts a*z - w*z into one fma, so an exact
        # zero on the CPU can come out as one rounding of a product here
        scale = max(1.0, float(np.abs(wh[k]).max()))
        np.testing.assert_allclose(out[k, 0], ref[0], rtol=1e-6, atol=1e-14 * scale * n)
        np.testing.assert_allclose(out[k, 1], ref[1], rtol=1e-6, atol=1e-14 * n)
        assert out[k, 2] == pytest.approx(ref[2], rel=1e-12)
        assert int(out[k, 3]) == int(ref[3]), (k, out[k], ref[:4])
    assert int(out[0, 3]) == 0
    if n > 1:
        assert int(out[1, 3]) & 1 and int(out[2, 3]) & 2


def test_check_reference_entry_reports(capfd):
    """eigen_GPU_check_FP writes the reference's report over each a_k and
    prints the failures."""
    n, nm, L = 16, 16, 4
    a0 = wl.sym_rand(L, n, nm, dtype=np.float32, seed=3)
    a_full = full_sym(a0)
    t = torch.from_numpy(a0.copy()).to(DEV)
    z, w, _ = eg.eigen_batch(t)
    w[3, 5] += 0.5
    af = torch.from_numpy(a_full).to(DEV)
    lib = eg.lib()
    lib.eigen_GPU_check_FP(L, nm, n, n, af.data_ptr(), w.data_ptr(), z.data_ptr(),
                           torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rep = af.cpu().numpy().reshape(L, -1)[:, :4]
    assert rep[0, 0] == 0 and rep[0, 1] == 0 and rep[3, 0] > 0
    assert rep[0, 3] == np.float32(2.0 ** -23)
    printed = capfd.readouterr().out
    assert "[000003] Accuracy error" in printed and "[000000]" not in printed
