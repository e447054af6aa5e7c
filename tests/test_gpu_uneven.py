"""GPU parity of the uneven-batch entry points (eigen_GPU_batch_uneven_FP/_DP).

Each matrix of a mixed-order batch must come out exactly as the oracle solves
it alone: same bar as the uniform path (eigenvalues within the reference
checker's TOL, residual and orthogonality flags clear, eigenvectors of
separated eigenvalues equal up to sign).  Layout variants (packed, explicit
offsets with gaps), out-of-range orders, both QL variants and a HIP-graph
replay are covered.
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import workloads as wl
from test_gpu_parity import check_batch

pytestmark = pytest.mark.gpu

DEV = "cuda"


def random_orders(L, lo, hi, seed):
    return np.random.default_rng(seed).integers(lo, hi + 1, size=L)


def make_mats(ns, dtype, seed):
    out = []
    for k, n in enumerate(ns):
        a = wl.sym_rand(1, int(n), int(n), dtype=dtype, seed=seed * 100003 + k)[0]  # (n, nm) col-major view
        out.append(np.ascontiguousarray(a.T))  # row r col c
    return out


def check_against_oracle(mats, res, dtype, solver=eg.IMTQL2):
    """Group the matrices by order, solve each group with the oracle, compare."""
    by_n = {}
    for k, m in enumerate(mats):
        by_n.setdefault(m.shape[0], []).append(k)
    osolver = oracle.TQL2 if solver == eg.TQL2 else oracle.IMTQL2
    for n, ks in sorted(by_n.items()):
        a0 = np.stack([mats[k].T for k in ks]).astype(dtype)  # (Lg, n, nm=n), a0[k, c, r]
        zo, wo, info_o = oracle.eig_batch(a0.copy(), osolver)
        assert info_o.max() == 0
        z = np.stack([res[k][0].T for k in ks])
        w = np.stack([res[k][1] for k in ks])
        check_batch(a0, z, w, dtype, wo, zo)


@pytest.mark.parametrize("solver", [eg.IMTQL2, eg.TQL2])
@pytest.mark.parametrize("dtype,lo,hi,nmax", [(np.float32, 1, 160, 160), (np.float64, 1, 100, 100)])
def test_uneven_packed_vs_oracle(dtype, lo, hi, nmax, solver):
    print(f"[uneven] packed dtype={np.dtype(dtype).name} n={lo}..{hi} solver={solver}", flush=True)
    ns = random_orders(240, lo, hi, seed=11 + solver)
    mats = make_mats(ns, dtype, seed=7)
    a, nv = eg.pack_uneven([torch.from_numpy(m) for m in mats], DEV)
    a, w, info = eg.eigen_batch_uneven(a, nv, nmax, solver=solver)
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    res = eg.unpack_uneven(a, w, [int(n) for n in ns])
    check_against_oracle(mats, res, dtype, solver)


@pytest.mark.parametrize("route", ["auto", "split"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uneven_offsets_gaps_and_bad_orders(dtype, route):
    """Explicit offsets in reverse order with NaN gaps; orders 0 and > nmax are
    flagged -3 and their storage is left untouched (also through the split
    route of the orders beyond 64)."""
    print(f"[uneven] offsets dtype={np.dtype(dtype).name} route={route}", flush=True)
    nmax = 70
    ns = list(random_orders(60, 1, nmax, seed=3))
    ns[5] = 0
    ns[17] = nmax + 9
    mats = make_mats([max(n, 1) for n in ns], dtype, seed=9)
    gap = 5
    aoff, woff, pa, pw = [0] * len(ns), [0] * len(ns), 0, 0
    for k in reversed(range(len(ns))):
        aoff[k], woff[k] = pa, pw
        sz = max(ns[k], 1)
        pa += sz * sz + gap
        pw += sz + gap
    a_h = np.full(pa, np.nan, dtype=dtype)
    w_h = np.full(pw, np.nan, dtype=dtype)
    for k, n in enumerate(ns):
        sz = max(n, 1)
        a_h[aoff[k]:aoff[k] + sz * sz] = mats[k].T.reshape(-1)
    a = torch.from_numpy(a_h.copy()).to(DEV)
    w = torch.from_numpy(w_h.copy()).to(DEV)
    nv = torch.tensor(ns, dtype=torch.int32, device=DEV)
    ao = torch.tensor(aoff, dtype=torch.int64, device=DEV)
    wo = torch.tensor(woff, dtype=torch.int64, device=DEV)
    solver = eg.IMTQL2 | (eg.FLAG_SPLIT if route == "split" else 0)
    a, w, info = eg.eigen_batch_uneven(a, nv, nmax, w=w, aoff=ao, woff=wo, solver=solver)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    a_o, w_o = a.cpu().numpy(), w.cpu().numpy()
    bad = {5, 17}
    for k in bad:
        assert info[k] == eg.UNEVEN_BAD_ORDER
        sz = max(ns[k], 1)
        np.testing.assert_array_equal(a_o[aoff[k]:aoff[k] + sz * sz], a_h[aoff[k]:aoff[k] + sz * sz])
        assert np.isnan(w_o[woff[k]:woff[k] + sz]).all()
    good = [k for k in range(len(ns)) if k not in bad]
    assert (info[good] == 0).all()
    # gaps untouched
    used = np.zeros(pa, bool)
    for k, n in enumerate(ns):
        sz = max(n, 1)
        used[aoff[k]:aoff[k] + sz * sz] = True
    assert np.isnan(a_o[~used]).all()
    res = {k: (a_o[aoff[k]:aoff[k] + ns[k] ** 2].reshape(ns[k], ns[k]).T, w_o[woff[k]:woff[k] + ns[k]])
           for k in good}
    check_against_oracle([mats[k] for k in good], [res[k] for k in good], dtype)


def test_uneven_matches_uniform_path():
    """A batch of one order through the uneven entry point agrees with the
    uniform entry point (same per-matrix solver)."""
    n, L = 48, 200
    a0 = wl.sym_rand(L, n, n, dtype=np.float32, seed=21)
    t = torch.from_numpy(a0.copy()).to(DEV)
    _, w1, _ = eg.eigen_batch(t, path="fused")
    a = torch.from_numpy(a0.copy()).to(DEV).reshape(-1)
    nv = torch.full((L,), n, dtype=torch.int32, device=DEV)
    a, w2, info = eg.eigen_batch_uneven(a, nv, n)
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    w1 = w1.cpu().numpy()
    w2 = w2.cpu().numpy().reshape(L, n)
    scale = np.abs(w1).max(axis=1, keepdims=True)
    assert np.max(np.abs(w1 - w2) / scale) <= 8 * np.finfo(np.float32).eps
    z2 = a.cpu().numpy().reshape(L, n, n)
    check_batch(a0, z2, w2, np.float32, w1, t.cpu().numpy())


def test_uneven_graph_replay():
    """The whole uneven call (binning + per-class launches) captured in a HIP
    graph: every replay equals the eager result."""
    ns = random_orders(500, 4, 128, seed=5)
    mats = make_mats(ns, np.float32, seed=13)
    src, nv = eg.pack_uneven([torch.from_numpy(m) for m in mats], DEV)
    nmax = 128
    wlen = int(ns.sum())
    a_e = src.clone()
    a_e, w_e, info_e = eg.eigen_batch_uneven(a_e, nv, nmax)
    torch.cuda.synchronize()
    assert int(info_e.abs().max().item()) == 0

    work = torch.empty_like(src)
    w = torch.empty(wlen, dtype=src.dtype, device=DEV)
    info = torch.empty(len(ns), dtype=torch.int32, device=DEV)
    ws = torch.empty(eg.uneven_bufsize(len(ns), nmax, 4), dtype=torch.uint8, device=DEV)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up on the capture stream
        work.copy_(src)
        eg.eigen_batch_uneven(work, nv, nmax, w=w, info=info, workspace=ws)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        work.copy_(src)
        eg.eigen_batch_uneven(work, nv, nmax, w=w, info=info, workspace=ws)
    for rep in range(3):
        work.fill_(0.0)
        w.fill_(0.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(work, a_e), f"replay {rep}: eigenvectors differ"
        assert torch.equal(w, w_e), f"replay {rep}: eigenvalues differ"
        assert int(info.abs().max().item()) == 0


@pytest.mark.parametrize("solver", [eg.IMTQL2, eg.TQL2])
@pytest.mark.parametrize("dtype,lo,hi,nmax", [(np.float32, 1, 128, 128), (np.float32, 60, 128, 140),
                                               (np.float64, 1, 256, 256), (np.float64, 50, 200, 300)])
@pytest.mark.parametrize("tiny", [False, True])
def test_uneven_split_route_vs_oracle(dtype, lo, hi, nmax, solver, tiny):
    """The split route of the class beyond order 64 (large-order split kernels
    driven by the sorted list), forced for a small batch; with the rotation log
    shrunk, every matrix of that class takes the route's workgroup fallback."""
    print(f"[uneven] split route dtype={np.dtype(dtype).name} n={lo}..{hi} nmax={nmax} solver={solver} "
          f"tiny={tiny}", flush=True)
    ns = random_orders(90, lo, hi, seed=17 + solver + hi)
    mats = make_mats(ns, dtype, seed=19)
    a, nv = eg.pack_uneven([torch.from_numpy(m) for m in mats], DEV)
    flags = eg.FLAG_SPLIT | (eg.FLAG_TINYLOG if tiny else 0)
    a, w, info = eg.eigen_batch_uneven(a, nv, nmax, solver=solver | flags)
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    res = eg.unpack_uneven(a, w, [int(n) for n in ns])
    check_against_oracle(mats, res, dtype, solver)


def test_uneven_split_route_by_count():
    """Without forcing: a class beyond order 64 with at least 4096 matrices
    (FP32) takes the split route on its own; every matrix passes the checker
    and a sample matches the oracle."""
    ns = np.concatenate([random_orders(4200, 65, 128, seed=31), random_orders(300, 2, 64, seed=32)])
    ns = np.random.default_rng(33).permutation(ns)
    mats = make_mats(ns, np.float32, seed=34)
    a0, nv = eg.pack_uneven([torch.from_numpy(m) for m in mats], DEV)
    a, w, info = eg.eigen_batch_uneven(a0.clone(), nv, 128)
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    res = eg.unpack_uneven(a, w, [int(n) for n in ns])
    pick = np.random.default_rng(35).choice(len(ns), size=160, replace=False)
    check_against_oracle([mats[k] for k in pick], [res[k] for k in pick], np.float32)
    # the fused kernels give the same eigenvalues
    a2, w2, info2 = eg.eigen_batch_uneven(a0.clone(), nv, 128, solver=eg.IMTQL2 | eg.FLAG_FUSED)
    torch.cuda.synchronize()
    assert int(info2.abs().max().item()) == 0
    wa, wb = w.cpu().numpy(), w2.cpu().numpy()
    off = np.concatenate([[0], np.cumsum(ns)])
    for k in range(len(ns)):
        x, y = wa[off[k]:off[k + 1]], wb[off[k]:off[k + 1]]
        assert np.max(np.abs(x - y)) <= 64 * np.finfo(np.float32).eps * max(np.abs(y).max(), 1e-30)


@pytest.mark.parametrize("dtype,count", [(np.float32, 4300), (np.float64, 1100)])
def test_uneven_split_route_by_count_order64(dtype, count):
    """Without forcing, with nmax > 64: a class (32, 64] of at least 4096 (FP32) /
    1024 (FP64) matrices takes the n <= 64 split kernels in list form, while the
    small class beyond 64 keeps the fused kernel; a sample matches the oracle
    and every matrix's eigenvalues match the fused kernels'."""
    print(f"[uneven] by count, class (32,64] dtype={np.dtype(dtype).name} count={count}", flush=True)
    ns = np.concatenate([random_orders(count, 33, 64, seed=41), random_orders(60, 65, 100, seed=42)])
    ns = np.random.default_rng(43).permutation(ns)
    mats = make_mats(ns, dtype, seed=44)
    a0, nv = eg.pack_uneven([torch.from_numpy(m) for m in mats], DEV)
    a, w, info = eg.eigen_batch_uneven(a0.clone(), nv, 100)
    torch.cuda.synchronize()
    assert int(info.abs().max().item()) == 0
    res = eg.unpack_uneven(a, w, [int(n) for n in ns])
    pick = np.random.default_rng(45).choice(len(ns), size=120, replace=False)
    check_against_oracle([mats[k] for k in pick], [res[k] for k in pick], dtype)
    a2, w2, info2 = eg.eigen_batch_uneven(a0.clone(), nv, 100, solver=eg.IMTQL2 | eg.FLAG_FUSED)
    torch.cuda.synchronize()
    assert int(info2.abs().max().item()) == 0
    wa, wb = w.cpu().numpy(), w2.cpu().numpy()
    eps = np.finfo(dtype).eps
    off = np.concatenate([[0], np.cumsum(ns)])
    for k in range(len(ns)):
        x, y = wa[off[k]:off[k + 1]], wb[off[k]:off[k + 1]]
        assert np.max(np.abs(x - y)) <= 64 * eps * max(np.abs(y).max(), 1e-30)


def test_uneven_invalid_args():
    a = torch.zeros(16, device=DEV)
    nv = torch.full((1,), 4, dtype=torch.int32, device=DEV)
    with pytest.raises(RuntimeError):
        eg.eigen_batch_uneven(a, nv, 0)
    with pytest.raises(RuntimeError):
        eg.eigen_batch_uneven(a, nv, 513)
    with pytest.raises(RuntimeError):
        eg.eigen_batch_uneven(a, nv, 4, solver=eg.JACOBI)
