#!/usr/bin/env python3
"""Benchmark of the batched symmetric eigensolver on one MI355X.

Workload (one "step" = one eigen_GPU_batch_FP call over the whole batch):
  L = 16384 LETKF ensemble-space matrices, n = nm = m = 64, FP32,
  A = ((k-1)/rho) I + Y^T R^-1 Y with k = 64 members, p = 400 observations,
  Y_ij ~ N(0, 0.1^2), R = I, rho = 1.1 (synthetic, seeded, generated on the GPU).

Timing follows the reference harness (main.cpp:259-298): before each step the
input is restored by a device-to-device copy outside the timed interval; each
step is bracketed by HIP events on the stream the library launches on; W
untimed warm-up steps precede K timed ones.  `value` = matrices per second over
the K timed solves (inputs resident in HBM).

Also reported: the roofline of the solver kernel (algorithmic flops (10/3) n^3
and algorithmic bytes 4 (n(n+1)/2 + n^2 + n) per matrix over the measured kernel
time), accuracy of the GPU eigenpairs on a sample (oracle checker), and the CPU
baseline: the C oracle (EISPACK restatement) on the same inputs with all host
threads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "eigeng-batched_amd"), os.path.join(ROOT, "oracle")]

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def make_letkf_gpu(L, k, p, rho, sigma, seed, dtype):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.empty((L, k, k), dtype=torch.float32, device="cuda")
    chunk = 4096
    for s in range(0, L, chunk):
        e = min(L, s + chunk)
        Y = torch.randn((e - s, p, k), generator=g, device="cuda", dtype=torch.float32) * sigma
        Yt = Y.transpose(1, 2).contiguous()
        A[s:e] = torch.bmm(Yt, Y)
    A += torch.eye(k, device="cuda", dtype=torch.float32) * ((k - 1) / rho)
    A = 0.5 * (A + A.transpose(1, 2))
    return A.to(dtype).contiguous()


def householder_q(G):
    """Q of the QR factorisation of each matrix of G (B, n, n), by Householder
    reflections written with batched torch ops (no solver library, so the
    generator also runs under hardware-counter profiling)."""
    import torch
    B, n, _ = G.shape
    A = G.clone()
    vs = []
    for k in range(n - 1):
        x = A[:, k:, k]
        alpha = torch.linalg.vector_norm(x, dim=1)
        sgn = torch.where(x[:, 0] >= 0, 1.0, -1.0).to(A.dtype)
        v = x.clone()
        v[:, 0] += sgn * alpha
        v = v / torch.linalg.vector_norm(v, dim=1, keepdim=True).clamp_min(1e-300)
        sub = A[:, k:, k:]
        sub -= 2.0 * v.unsqueeze(2) * torch.bmm(v.unsqueeze(1), sub)
        vs.append(v)
    Q = torch.eye(n, dtype=A.dtype, device=A.device).expand(B, n, n).contiguous()
    for k in reversed(range(n - 1)):
        v = vs[k]
        sub = Q[:, k:, :]
        sub -= 2.0 * v.unsqueeze(2) * torch.bmm(v.unsqueeze(1), sub)
    return Q


def make_qdq_gpu(L, n, seed, dtype, frac=0.05, gap=1e-8):
    """A = Q diag(lam) Q^T, Q from the QR of a Gaussian matrix, lam ~ U[-1,1]
    with a fraction `frac` of the eigenvalues replaced by pairs split by `gap`."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.empty((L, n, n), dtype=torch.float64, device="cuda")
    npairs = max(1, int(round(frac * n / 2)))
    chunk = 8192                    # few launches: the generator also runs under profiling
    for s in range(0, L, chunk):
        e = min(L, s + chunk)
        G = torch.randn((e - s, n, n), generator=g, device="cuda", dtype=torch.float64)
        Q = householder_q(G)
        lam = torch.rand((e - s, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
        idx = torch.argsort(torch.rand((e - s, n), generator=g, device="cuda"), dim=1)[:, : 2 * npairs]
        src = torch.gather(lam, 1, idx[:, 0::2])
        lam.scatter_(1, idx[:, 1::2], src + gap)
        QL = Q * lam[:, None, :]
        A[s:e] = torch.bmm(QL, Q.transpose(1, 2).contiguous())
    A = 0.5 * (A + A.transpose(1, 2))
    return A.to(dtype).contiguous()


def time_solver(a0, solver, steps, warmup):
    """Mean HIP-event time (ms) of one eigen_GPU_batch call over the batch."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    L, n, _ = a0.shape
    a = torch.empty_like(a0)
    w = torch.empty((L, n), dtype=a0.dtype, device="cuda")
    ws = torch.empty(max(eg.bufsize(L, n, n, n, a.element_size(), solver), 1), dtype=torch.uint8,
                     device="cuda")
    info = torch.empty(L, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        a.copy_(a0)
        eg.eigen_batch(a, w=w, solver=solver, info=info, workspace=ws, stream=stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    wall0 = time.perf_counter()
    for i in range(steps):
        a.copy_(a0)
        ev[i][0].record(stream)
        eg.eigen_batch(a, w=w, solver=solver, info=info, workspace=ws, stream=stream)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = float(np.mean([s.elapsed_time(e) for s, e in ev]))
    return ms, wall, a, w, info


def tuned_line(a0, plain_ms, steps, warmup):
    """The autotuner on the batch (every candidate path and launch shape timed
    once on these inputs), then the remembered winner timed like the plain
    call, with the GPU checker on its results (equal-residual comparison)."""
    import torch
    import eigeng_batched as eg
    best, best_ms, table, _, _ = eg.autotune(a0)
    ms, _, a, w, info = time_solver(a0, 0, steps, warmup)
    acc = accuracy_all(a0, a, w)
    bad = int((info != 0).sum().item())
    L = a0.shape[0]
    del a, w
    torch.cuda.empty_cache()
    eg.tune_reset()
    return {"value": L / (ms * 1e-3), "unit": "matrices/s", "ms_per_step": ms, "best_bits": hex(best),
            "candidates_ms": {hex(k): v for k, v in table.items()}, "vs_plain": plain_ms / ms,
            "nonconverged": bad, "accuracy_all": acc}


def time_phases(a0, steps, warmup, path="split"):
    """Per-phase HIP-event times (ms) of the split paths: the phases are issued
    as separate library calls on one stream with events between them (n <= 64:
    tri = eig_tri_kernel, ql = eig_ql_reg_kernel, apply = eig_apply_kernel plus
    the overflow-fallback launch; larger n: eig_tri_reg_kernel (n <= 256),
    eig_ql_lane_kernel, eig_apply_large_kernel plus the fallback launch, whose
    workgroups exit at once when no matrix overflowed).  Returns {phase: mean ms}."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    L, n, _ = a0.shape
    a = torch.empty_like(a0)
    w = torch.empty((L, n), dtype=a0.dtype, device="cuda")
    ws = torch.empty(max(eg.bufsize(L, n, n, n, a.element_size(), eg.PATHS[path]), 1), dtype=torch.uint8,
                     device="cuda")
    info = torch.empty(L, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    names = ["tri", "ql", "apply"]
    rec = {p: [] for p in names}
    for i in range(warmup + steps):
        a.copy_(a0)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        for j, p in enumerate(names):
            eg.eigen_batch(a, w=w, info=info, workspace=ws, stream=stream, path=path, phases=[p])
            ev[j + 1].record(stream)
        torch.cuda.synchronize()
        if i >= warmup:
            for j, p in enumerate(names):
                rec[p].append(ev[j].elapsed_time(ev[j + 1]))
    return {p: float(np.mean(v)) for p, v in rec.items()}, a, w, info


def make_uneven_gpu(count, nlo, nhi, seed, lam_hi=1.0, lam_lo=1e-6):
    """Uneven batch: `count` matrices with orders ~ U{nlo..nhi} in random order,
    A = Q diag(lam) Q^T with Q from the QR of a Gaussian matrix and lam
    log-spaced over [lam_lo, lam_hi] in random order.  Returns the packed FP32
    batch (matrix k column-major after matrix k-1), the orders, and the same
    matrices grouped by order ({n: (c, n, n) tensor}) for one call per order."""
    import numpy as np
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    ns = torch.randint(nlo, nhi + 1, (count,), generator=g, device="cuda", dtype=torch.int32)
    ns_h = ns.cpu().numpy().astype(np.int64)
    sizes = ns_h * ns_h
    aoff = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    packed = torch.empty(int(sizes.sum()), dtype=torch.float32, device="cuda")
    grouped = {}
    for n in range(nlo, nhi + 1):
        idx = np.nonzero(ns_h == n)[0]
        c = len(idx)
        if c == 0:
            continue
        lam0 = torch.logspace(np.log10(lam_hi), np.log10(lam_lo), n, dtype=torch.float64, device="cuda")
        A = torch.empty((c, n, n), dtype=torch.float32, device="cuda")
        for s0 in range(0, c, 256):
            e0 = min(c, s0 + 256)
            G = torch.randn((e0 - s0, n, n), generator=g, device="cuda", dtype=torch.float64)
            Q = householder_q(G)
            perm = torch.argsort(torch.rand((e0 - s0, n), generator=g, device="cuda"), dim=1)
            lam = lam0[perm]
            M = torch.bmm(Q * lam[:, None, :], Q.transpose(1, 2).contiguous())
            A[s0:e0] = (0.5 * (M + M.transpose(1, 2))).to(torch.float32)
        grouped[n] = A
        dst = (torch.from_numpy(aoff[idx]).to("cuda")[:, None]
               + torch.arange(n * n, device="cuda")[None, :]).reshape(-1)
        packed.index_copy_(0, dst, A.reshape(-1))
    return packed, ns, grouped


def bench_uneven(count, nlo, nhi, steps, warmup, seed=0xE16E0003):
    """Uneven batch, one size-binned call vs one uniform call per order."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    import oracle
    packed, nv, grouped = make_uneven_gpu(count, nlo, nhi, seed)
    ns_h = nv.cpu().numpy().astype(np.int64)
    stream = torch.cuda.current_stream()
    work = torch.empty_like(packed)
    w = torch.empty(int(ns_h.sum()), dtype=torch.float32, device="cuda")
    info = torch.empty(count, dtype=torch.int32, device="cuda")
    ws = torch.empty(eg.uneven_bufsize(count, nhi, 4), dtype=torch.uint8, device="cuda")
    t_bin = []
    for i in range(warmup + steps):
        work.copy_(packed)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        eg.eigen_batch_uneven(work, nv, nhi, w=w, info=info, workspace=ws, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if i >= warmup:
            t_bin.append(e0.elapsed_time(e1))
    bad = int((info != 0).sum().item())
    # accuracy on a sample
    work_h = None
    rng = np.random.default_rng(1)
    sample = rng.choice(count, size=min(128, count), replace=False)
    aoff = np.concatenate([[0], np.cumsum(ns_h * ns_h)[:-1]])
    woff = np.concatenate([[0], np.cumsum(ns_h)[:-1]])
    worst_rel, worst_orth, fails = 0.0, 0.0, 0
    for k in sample:
        n = int(ns_h[k])
        a_in = packed[aoff[k]:aoff[k] + n * n].reshape(n, n).cpu().numpy()
        z = work[aoff[k]:aoff[k] + n * n].reshape(n, n).cpu().numpy()
        wk = w[woff[k]:woff[k] + n].cpu().numpy()
        out = oracle.check(a_in.copy(), wk, z)
        fails += int(out[3] != 0)
        worst_rel = max(worst_rel, float(out[5]))
        worst_orth = max(worst_orth, float(out[1]))
    del work, work_h
    # one uniform call per order (matrices already grouped by order)
    works = {n: torch.empty_like(A) for n, A in grouped.items()}
    wsz = max(eg.bufsize(A.shape[0], n, n, n, 4) for n, A in grouped.items())
    ws2 = torch.empty(wsz, dtype=torch.uint8, device="cuda")
    wn = {n: torch.empty((A.shape[0], n), dtype=torch.float32, device="cuda") for n, A in grouped.items()}
    t_per = []
    for i in range(warmup + steps):
        for n, A in grouped.items():
            works[n].copy_(A)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for n in grouped:
            eg.eigen_batch(works[n], w=wn[n], workspace=ws2, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        if i >= warmup:
            t_per.append(e0.elapsed_time(e1))
    ms_bin, ms_per = float(np.mean(t_bin)), float(np.mean(t_per))
    mean_b = float(np.mean(4 * (ns_h * (ns_h + 1) / 2 + ns_h * ns_h + ns_h)))
    mean_f = float(np.mean((10.0 / 3.0) * ns_h.astype(np.float64) ** 3))
    return {"workload": f"{count} matrices, n ~ U{{{nlo}..{nhi}}} (nm = n), FP32, A = Q diag(lam) Q^T, "
                        f"Q = qr(Gaussian), lam log-spaced 1..1e-6 shuffled, packed in random order",
            "value": count / (ms_bin * 1e-3), "unit": "matrices/s", "ms_per_step": ms_bin,
            "one_call_per_order": {"calls": len(grouped), "ms_per_step": ms_per,
                                   "value": count / (ms_per * 1e-3)},
            "speedup_vs_per_order": ms_per / ms_bin,
            "achieved_tflops": count * mean_f / (ms_bin * 1e-3) / 1e12,
            "achieved_gbs": count * mean_b / (ms_bin * 1e-3) / 1e9,
            "nonconverged": bad,
            "accuracy": {"sample": len(sample), "checker_failures": fails,
                         "max_rel_residual_F": worst_rel, "max_orth": worst_orth}}


def bench_n32(steps, warmup, tuned=True):
    """L = 1024, n = nm = m = 32, FP64, A = (B + B^T)/2 with B ~ U[-1,1] (seeded,
    generated on the GPU): time, HBM and VALU roofs of the solver kernel, every
    matrix checked on the GPU and compared with the CPU oracle matrix by matrix."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    import oracle
    L, n = 1024, 32
    g = torch.Generator(device="cuda")
    g.manual_seed(0xE16E0001)
    B = torch.rand((L, n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    a0 = (0.5 * (B + B.transpose(1, 2))).contiguous()
    ms, _, a, w, info = time_solver(a0, 0, steps, warmup)
    acc = accuracy_all(a0, a, w)
    _, wo, io = oracle.eig_batch(a0.cpu().numpy(), oracle.IMTQL2)
    wg = w.cpu().numpy()
    scale = np.abs(wo).max(axis=1, keepdims=True)
    par = float(np.max(np.abs(wg - wo) / scale))
    p = eg.launch_params(L, n, 8)
    tl = tuned_line(a0, ms, steps, warmup) if tuned else None
    b = 8 * (n * (n + 1) / 2 + n * n + n)
    f = (10.0 / 3.0) * n ** 3
    gbs = L * b / (ms * 1e-3) / 1e9
    tfs = L * f / (ms * 1e-3) / 1e12
    del a0, a, w, B
    torch.cuda.empty_cache()
    return {"workload": "L=1024 n=nm=m=32 FP64, A = (B + B^T)/2, B ~ U[-1,1] (seeded)",
            "value": L / (ms * 1e-3), "unit": "matrices/s", "ms_per_step": ms,
            "nonconverged": int((info != 0).sum().item()) + int(io.max()),
            "accuracy_all": acc, "oracle_max_eig_err_rel": par,
            "launch": {k: p[k] for k in ("grpSZ", "numTB", "numGR", "numTH")},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": gbs / PEAK_HBM_GBS, "bytes_per_matrix": b,
                         "kernel": "split path: eig_tri_kernel<double,32> + eig_ql_reg_kernel + eig_apply_kernel",
                         "valu_tflops": tfs, "valu_frac_fp64": tfs / PEAK_FP64_TFLOPS},
            "tuned": tl}


def bench_letkf48(steps, warmup):
    """The learned-warm-start workload's shape solved by the plain solver: one
    time step of L = 65536 LETKF matrices, n = 48, FP32, k = 48 members,
    p = 300 observations, rho = 1.05, Y ~ N(0, 0.1^2)."""
    import torch
    L, n = 65536, 48
    a0 = make_letkf_gpu(L, n, 300, 1.05, 0.1, seed=0xE16E0005, dtype=torch.float32)
    ms, _, a, w, info = time_solver(a0, 0, steps, warmup)
    acc = accuracy_all(a0, a, w)
    bad = int((info != 0).sum().item())
    del a0, a, w
    torch.cuda.empty_cache()
    return {"workload": "one step of the drifting-ensemble stream: L=65536 LETKF matrices n=48 FP32, "
                        "k=48 p=300 rho=1.05 sigma=0.1, plain solver (no learned warm start)",
            "value": L / (ms * 1e-3), "unit": "matrices/s", "ms_per_step": ms, "nonconverged": bad,
            "accuracy_all": acc,
            "achieved_tflops": L * (10.0 / 3.0) * n ** 3 / (ms * 1e-3) / 1e12}


def bench_letkf_step(steps, warmup, L=16384, k=64, p=400, rho=1.1, sigma=0.1, seed=0xE16E0012):
    """The whole LETKF analysis step on the GPU for L grid points: ingest
    A = (k-1)/rho I + Y^T Y (gram, MFMA), eigenpairs (the solver), consumer
    T = sqrt(k-1) A^-1/2 (matpow, MFMA).  Y ~ N(0, sigma^2), p x k per point."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    y = torch.randn((L, k, p), generator=g, device="cuda", dtype=torch.float32) * sigma   # y[i, c, o] = Y(o, c)
    a = torch.empty((L, k, k), dtype=torch.float32, device="cuda")
    t = torch.empty_like(a)
    w = torch.empty((L, k), dtype=torch.float32, device="cuda")
    info = torch.empty(L, dtype=torch.int32, device="cuda")
    ws = torch.empty(eg.bufsize(L, k, k, k, 4), dtype=torch.uint8, device="cuda")
    alpha, scale = (k - 1) / rho, float(np.sqrt(k - 1))

    def step():
        eg.gram(y, p, alpha, a=a)
        e1.record()
        eg.eigen_batch(a, w=w, info=info, workspace=ws)
        e2.record()
        eg.matpow(a, w, -0.5, scale, c=t)

    for _ in range(warmup):
        e1, e2 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        step()
    torch.cuda.synchronize()
    tot = {"ingest": 0.0, "solve": 0.0, "consumer": 0.0}
    for _ in range(steps):
        e0, e1, e2, e3 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
        e0.record()
        step()
        e3.record()
        torch.cuda.synchronize()
        tot["ingest"] += e0.elapsed_time(e1)
        tot["solve"] += e1.elapsed_time(e2)
        tot["consumer"] += e2.elapsed_time(e3)
    ph = {kk: v / steps for kk, v in tot.items()}
    ms = sum(ph.values())
    # accuracy of T on a sample against FP64 numpy
    idx = np.linspace(0, L - 1, 64).astype(int)
    yh = y[idx].cpu().numpy().astype(np.float64)
    th = t[idx].cpu().numpy().astype(np.float64)
    worst = 0.0
    for j in range(len(idx)):
        Y = yh[j].T
        A = alpha * np.eye(k) + Y.T @ Y
        lam, V = np.linalg.eigh(A)
        Tref = scale * (V / np.sqrt(lam)) @ V.T
        worst = max(worst, float(np.linalg.norm(th[j].T - Tref) / np.linalg.norm(Tref)))
    bad = int((info != 0).sum().item())
    tri = k * (k + 1) / 2
    ing_bytes = 4.0 * (p * k + k * k)          # read Y, write A (both triangles)
    ing_flops = 2.0 * p * tri                  # the triangle of Y^T Y
    con_bytes = 4.0 * (k * k + k + k * k)      # read Z and w, write T
    con_flops = 2.0 * k * tri                  # the triangle of Z D Z^T
    del y, a, t, w, ws
    torch.cuda.empty_cache()

    measured = {}
    tfile = os.path.join(ROOT, "profiles", "traffic_letkf_step_n64_fp32.json")
    if (L, k, p) == (16384, 64, 400) and os.path.exists(tfile):
        with open(tfile) as f:
            measured = json.load(f)["per_kernel_bytes"]

    def roof(bytes_, flops, ms_, kernel):
        gbs = L * bytes_ / (ms_ * 1e-3) / 1e9
        tfs = L * flops / (ms_ * 1e-3) / 1e12
        return {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": gbs / PEAK_HBM_GBS, "traffic": measured.get(kernel), "kernel": kernel,
                "bytes_per_matrix": bytes_, "mfma_tflops": tfs, "mfma_frac_fp32": tfs / PEAK_FP32_TFLOPS}

    return {"workload": f"LETKF analysis step on the GPU: L={L} points, k={k} members, p={p} observations, "
                        f"rho={rho}, FP32: gram ingest + eigensolve + sqrt(k-1) A^-1/2 consumer",
            "value": L / (ms * 1e-3), "unit": "matrices/s", "ms_per_step": ms, "phase_ms": ph,
            "nonconverged": bad, "max_rel_err_T_vs_fp64": worst,
            "ingest_roofline": roof(ing_bytes, ing_flops, ph["ingest"], "eig_gram_wave_kernel<float,4,true>"),
            "consumer_roofline": roof(con_bytes, con_flops, ph["consumer"], "eig_symdmt_wave_kernel<float,4>")}


def vendor_baseline(a0, reps=2, max_seconds=3.0):
    """The vendor library on the same inputs: hipSOLVER's syevjBatched (Jacobi,
    rocSOLVER underneath), as the reference's own comparator does
    (hipsolver.hpp:13-143, main.cpp:417-430).  Timed with HIP events like the
    solver; its eigenpairs are checked with the same GPU checker."""
    import ctypes
    import torch
    import eigeng_batched as eg
    try:
        hs = ctypes.CDLL("libhipsolver.so")
    except OSError:
        try:
            hs = ctypes.CDLL("/opt/rocm/lib/libhipsolver.so")
        except OSError as e:
            return {"error": f"libhipsolver not loadable: {e}"}
    vp, ci = ctypes.c_void_p, ctypes.c_int
    L, n, nm = a0.shape
    tag = "S" if a0.dtype == torch.float32 else "D"
    f = getattr(hs, f"hipsolver{tag}syevjBatched")
    fb = getattr(hs, f"hipsolver{tag}syevjBatched_bufferSize")
    f.argtypes = [vp, ci, ci, ci, vp, ci, vp, vp, ci, vp, vp, ci]
    fb.argtypes = [vp, ci, ci, ci, vp, ci, vp, ctypes.POINTER(ci), vp, ci]
    f.restype = fb.restype = ci
    handle, params = vp(), vp()
    if hs.hipsolverCreate(ctypes.byref(handle)) != 0 or hs.hipsolverCreateSyevjInfo(ctypes.byref(params)) != 0:
        return {"error": "hipsolverCreate failed"}
    stream = torch.cuda.current_stream()
    hs.hipsolverSetStream(handle, vp(stream.cuda_stream))
    EIG_MODE_VECTOR, FILL_UPPER = 202, 121
    a = a0.clone()
    w = torch.empty((L, n), dtype=a0.dtype, device="cuda")
    info = torch.zeros(L, dtype=torch.int32, device="cuda")
    lwork = ci(0)
    rc = fb(handle, EIG_MODE_VECTOR, FILL_UPPER, n, vp(a.data_ptr()), nm, vp(w.data_ptr()), ctypes.byref(lwork),
            params, L)
    work = torch.empty(max(lwork.value, 1), dtype=a0.dtype, device="cuda")
    best = None
    for _ in range(reps):
        a.copy_(a0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = f(handle, EIG_MODE_VECTOR, FILL_UPPER, n, vp(a.data_ptr()), nm, vp(w.data_ptr()), vp(work.data_ptr()),
               lwork.value, vp(info.data_ptr()), params, L)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
        if rc != 0 or ms > max_seconds * 1e3:
            break
    acc = accuracy_all(a0, a, w) if rc == 0 else None
    hs.hipsolverDestroySyevjInfo(params)
    hs.hipsolverDestroy(handle)
    return {"library": f"hipSOLVER hipsolver{tag}syevjBatched (default tolerance and sweeps)", "status": rc,
            "value": L / (best * 1e-3), "unit": "matrices/s", "ms_per_call": best, "matrices": L,
            "nonconverged": int((info != 0).sum().item()), "accuracy_all": acc}


def _traffic(name):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy
    of known size) from profiles/, or None."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def bench_n256(steps, warmup, vendor=True, tuned=True, mixed=True):
    """L = 8192, n = 256, FP64 line: whole-call time, per-phase times, roofline,
    accuracy of every matrix."""
    import torch
    L2, n2 = 8192, 256
    q0 = make_qdq_gpu(L2, n2, seed=0xE16E0004, dtype=torch.float64)
    ms2, _, a2, w2, info2 = time_solver(q0, 0, steps, warmup)
    acc2 = accuracy_all(q0, a2, w2)
    bad = int((info2 != 0).sum().item())
    del a2, w2
    torch.cuda.empty_cache()
    tuned2 = tuned_line(q0, ms2, steps, warmup) if tuned else None
    vend = vendor_baseline(q0[:512].contiguous(), reps=1) if vendor else None
    mixed = bench_mixed(q0, ms2, steps, warmup) if mixed else None
    chunks2 = chunk_tradeoff(q0, (0, 1), steps=steps, warmup=warmup) if mixed else None
    ph2, a2, w2, _ = time_phases(q0, 2, 1)
    del q0, a2, w2
    torch.cuda.empty_cache()
    kms2 = sum(ph2.values())
    f2 = (10.0 / 3.0) * n2 ** 3
    b2 = 8 * (n2 * (n2 + 1) / 2 + n2 * n2 + n2)
    return {"workload": "L=8192 n=256 FP64, A = Q diag(lam) Q^T, lam ~ U[-1,1], 5% pairs split by 1e-8",
            "value": L2 / (ms2 * 1e-3), "unit": "matrices/s", "ms_per_step": ms2, "nonconverged": bad,
            "accuracy_all": acc2, "tuned": tuned2, "vendor_baseline_512": vend, "mixed": mixed,
            "chunk_tradeoff": chunks2,
            "roofline": {"bound": "valu", "achieved": L2 * f2 / (kms2 * 1e-3) / 1e12,
                         "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                         "frac": L2 * f2 / (kms2 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                         "kernel": "eig_tri_reg_kernel + eig_ql_lane_kernel + eig_apply_large_kernel "
                                   "(+ overflow-fallback launch)",
                         "kernel_ms": kms2, "phase_ms": ph2,
                         "note": "(10/3)n^3 flops per matrix against the FP64 vector peak"},
            "roofline_hbm": {"achieved": L2 * b2 / (kms2 * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "bytes_per_matrix": b2,
                             "traffic": _traffic("traffic_eig_split_n256_fp64.json")}}


def gemm_f64_roofline(batch, n, reps=5):
    """HIP-event time of one FP64 matrix-core product of the refinement's shape
    (batch x (n x n) x (n x n), eig_gemm_kernel<GK_F64>), on the stream it is
    launched on: achieved TFLOP/s at 2 n^3 flops per product."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    x = torch.randn((batch, n, n), dtype=torch.float64, device="cuda")
    y = torch.randn((batch, n, n), dtype=torch.float64, device="cuda")
    c = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    eg.gemm(x, y, c=c, stream=stream)
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        eg.gemm(x, y, c=c, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    ms = float(np.mean(ts))
    tf = batch * 2.0 * n ** 3 / (ms * 1e-3) / 1e12
    del x, y, c
    return {"kernel": "eig_gemm_kernel<GK_F64> (v_mfma_f64_16x16x4f64)", "batch": batch, "n": n,
            "ms": ms, "achieved": tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_FP64_TFLOPS}


def bench_mixed(q0, plain_ms, steps, warmup):
    """The FP64 interface's mixed-precision solver (FP32 solve, FP64 refinement on
    the matrix cores, plain fallback) on the same batch: time, its ratio to the
    plain FP64 solve, every matrix checked on the GPU (relative residual
    ||AZ - ZD||_F / max|w| and orthogonality), and the FP32 solve alone."""
    import torch
    import eigeng_batched as eg
    L, n, _ = q0.shape
    ms, _, a, w, info = time_solver(q0, eg.MIXED, steps, warmup)
    acc = accuracy_all(q0, a, w)
    bad = int((info != 0).sum().item())
    del a, w
    torch.cuda.empty_cache()
    ms16, _, a, w, info = time_solver(q0, eg.MIXED | eg.FLAG_MIXED_WY16, steps, warmup)
    acc16 = accuracy_all(q0, a, w)
    bad16 = int((info != 0).sum().item())
    del a, w
    torch.cuda.empty_cache()
    q32 = q0.float().contiguous()
    ms32, _, a32, w32, _ = time_solver(q32, 0, steps, warmup)
    del q32, a32, w32
    torch.cuda.empty_cache()
    return {"value": L / (ms * 1e-3), "unit": "matrices/s", "ms_per_step": ms, "vs_plain": plain_ms / ms,
            "fp32_solve_ms": ms32, "refinement_ms": ms - ms32, "nonconverged": bad, "accuracy_all": acc,
            "relative_residual_max": acc["max_err1_rel"],
            "fp16_blocked_back_transform": {"ms_per_step": ms16, "value": L / (ms16 * 1e-3),
                                            "vs_plain": plain_ms / ms16, "nonconverged": bad16,
                                            "accuracy_all": acc16},
            "gemm_roofline": gemm_f64_roofline(2048, n)}


def timed(fn, reps=3):
    """Mean HIP-event time (ms) of fn() on the current stream (one untimed call first)."""
    import numpy as np
    import torch
    fn()
    s = torch.cuda.current_stream()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.mean(ts))


def bench_warmstart(L=65536, n=48, p=300, rho=1.05, hidden=512, train_steps=50, total_steps=200,
                    eval_steps=(50, 100, 150, 199), max_sweeps=6, batch=4096, lr=1e-3):
    """The learned warm start on the drifting stream: Y_t = Y_{t-1} + N(0, 0.01^2),
    Y_0 ~ N(0, 0.1^2), L matrices A_t = ((k-1)/rho) I + Y_t^T Y_t (k = n = 48,
    p = 300, rho = 1.05).  The network (3 layers, hidden 512) is trained by the
    library's kernels on the plain solver's eigenpairs of steps 0..train_steps-1
    (one epoch per step, minibatches of 4096, Adam); at the evaluation steps the
    plain solve, the network + warm solve, and the ablations (identity start,
    previous step's eigenvectors) are timed and checked."""
    import numpy as np
    import torch
    import eigeng_batched as eg
    import warmstart
    torch.manual_seed(0)
    stream = warmstart.DriftingStream(L, n, p=p, rho=rho)
    net = warmstart.WarmNet(n, hidden=hidden, batch=batch, seed=1)
    t0 = time.perf_counter()
    hist = warmstart.train_on_stream(net, stream, steps=train_steps, epochs=1, lr=lr)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    ws_warm = torch.empty(int(eg.lib().eigen_GPU_batch_warm_bufsize(L, n, 4)), dtype=torch.uint8, device="cuda")
    ws_plain = torch.empty(max(eg.bufsize(L, n, n, n, 4), 1), dtype=torch.uint8, device="cuda")
    eye = torch.eye(n, device="cuda").expand(L, n, n).contiguous()
    evals = []
    prev_v = None
    while stream.t < total_steps:
        t = stream.t
        need_prev = (t + 1) in eval_steps
        if t in eval_steps or need_prev:
            a = stream.matrices()
            z = torch.empty_like(a)
            w = torch.empty((L, n), device="cuda")
            info = torch.empty(L, dtype=torch.int32, device="cuda")

            def plain():
                z.copy_(a)
                eg.eigen_batch(z, w=w, info=info, workspace=ws_plain)
            if t in eval_steps:
                ms_plain = timed(plain)
                acc_plain = accuracy_all(a, z, w)
                q0 = torch.empty_like(a)
                sw = torch.empty(L, dtype=torch.int32, device="cuda")
                res = {"step": t, "plain_ms": ms_plain, "plain_accuracy": acc_plain}
                starts = {"network": None, "identity": eye}
                if prev_v is not None:
                    starts["previous_step"] = prev_v
                for name, q in starts.items():
                    def run(q=q):
                        z.copy_(a)
                        qq = net.forward(a, q0) if q is None else q
                        eg.eigen_batch_warm(z, qq, max_sweeps=max_sweeps, w=w, info=info, sweeps=sw,
                                            workspace=ws_warm)
                    ms = timed(run)
                    acc = accuracy_all(a, z, w)
                    swn = sw.cpu().numpy()
                    res[name] = {"ms": ms, "matrices_per_s": L / (ms * 1e-3), "vs_plain": ms_plain / ms,
                                 "mean_sweeps": float(swn.mean()),
                                 "fallbacks": int((swn >= max_sweeps).sum()), "accuracy_all": acc}
                res["network_forward_ms"] = timed(lambda: net.forward(a, q0))
                evals.append(res)
                print(f"[warmstart] step {t}: plain {ms_plain:.2f} ms, network {res['network']['ms']:.2f} ms "
                      f"({res['network']['mean_sweeps']:.2f} sweeps), identity {res['identity']['ms']:.2f} ms "
                      f"({res['identity']['mean_sweeps']:.2f} sweeps)", file=sys.stderr, flush=True)
            if need_prev:
                plain()
                prev_v = z.clone()
            del a, z
        stream.advance()
    del stream
    torch.cuda.empty_cache()
    mean = lambda k, f: float(np.mean([e[k][f] for e in evals if k in e]))
    return {"workload": f"drifting stream of L={L} LETKF matrices n={n} FP32, k={n} p={p} rho={rho}, "
                        f"Y_t = Y_(t-1) + N(0, 0.01^2), Y_0 ~ N(0, 0.1^2); network trained on steps "
                        f"0..{train_steps - 1}, evaluated at steps {list(eval_steps)}",
            "network": {"layers": [n * (n + 1) // 2, hidden, hidden, n * n], "params": net.count,
                        "train_seconds": train_s, "adam_steps": net.step,
                        "loss_first_last": [hist[0], hist[-1]]},
            "max_sweeps": max_sweeps,
            "value": L / (mean("network", "ms") * 1e-3), "unit": "matrices/s",
            "plain_value": L / (float(np.mean([e["plain_ms"] for e in evals])) * 1e-3),
            "vs_plain": {k: mean(k, "vs_plain") for k in ("network", "identity", "previous_step")},
            "mean_sweeps": {k: mean(k, "mean_sweeps") for k in ("network", "identity", "previous_step")},
            "evals": evals}


def chunk_tradeoff(a0, shifts, steps=5, warmup=1):
    """Time and workspace of the same solve with the split path's batch chunk
    divided by 2^shift (tuning bits 18-19): the memory / time trade-off."""
    import torch
    import eigeng_batched as eg
    L, n, nm = a0.shape
    out = {}
    for sh in shifts:
        bits = sh << eg.CHUNK_SHIFT
        ms, _, a, w, info = time_solver(a0, bits, steps, warmup)
        out[f"chunk/{1 << sh}"] = {"ms_per_step": ms, "value": L / (ms * 1e-3),
                                   "workspace_bytes": eg.bufsize(L, nm, n, n, a0.element_size(), bits),
                                   "nonconverged": int((info != 0).sum().item())}
        del a, w
        torch.cuda.empty_cache()
    return out


def bench_n128(steps, warmup):
    """L = 8192, n = 128, FP32 line (the register split path, the default for
    FP32 64 < n <= 128 from 2048 matrices): whole-call time, per-phase times,
    roofline, accuracy of every matrix."""
    import torch
    L2, n2 = 8192, 128
    q0 = make_qdq_gpu(L2, n2, seed=0xE16E0005, dtype=torch.float32)
    ms2, _, a2, w2, info2 = time_solver(q0, 0, steps, warmup)
    acc2 = accuracy_all(q0, a2, w2)
    bad = int((info2 != 0).sum().item())
    del a2, w2
    torch.cuda.empty_cache()
    ph2, a2, w2, _ = time_phases(q0, 2, 1, path="split-reg128")
    del q0, a2, w2
    torch.cuda.empty_cache()
    kms2 = sum(ph2.values())
    f2 = (10.0 / 3.0) * n2 ** 3
    b2 = 4 * (n2 * (n2 + 1) / 2 + n2 * n2 + n2)
    return {"workload": "L=8192 n=128 FP32, A = Q diag(lam) Q^T, lam ~ U[-1,1], 5% pairs split by 1e-8",
            "value": L2 / (ms2 * 1e-3), "unit": "matrices/s", "ms_per_step": ms2, "nonconverged": bad,
            "accuracy_all": acc2,
            "roofline": {"bound": "valu", "achieved": L2 * f2 / (kms2 * 1e-3) / 1e12,
                         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": L2 * f2 / (kms2 * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                         "kernel": "eig_tri_kernel<float,128> + eig_ql_reg_kernel<float,128> + eig_apply_kernel<float,128> "
                                   "(+ overflow-fallback launch)",
                         "kernel_ms": kms2, "phase_ms": ph2,
                         "note": "(10/3)n^3 flops per matrix against the FP32 vector peak"},
            "roofline_hbm": {"achieved": L2 * b2 / (kms2 * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "bytes_per_matrix": b2,
                             "traffic": _traffic("traffic_eig_split128_n128_fp32.json")}}


def host_threads():
    """Threads for the CPU legs: the CPUs this process may run on
    (os.sched_getaffinity), capped by OMP_NUM_THREADS when the host sets it
    (a shared GPU host gives each GPU's jobs a share of its cores that way)."""
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    share = int(env) if env.isdigit() and int(env) > 0 else visible
    return min(visible, share), visible


def time_cpu(a_host, solver, threads, min_seconds, max_reps=50):
    """Matrices/s of the C oracle over repeated passes of a_host (bounded time)."""
    import numpy as np
    import oracle
    reps, t0, total = 0, time.perf_counter(), 0
    while True:
        _, _, info = oracle.eig_batch(a_host, solver, threads=threads)
        reps += 1
        total += a_host.shape[0]
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            break
    return {"value": total / el, "matrices": int(total), "passes": reps, "seconds": el,
            "nonconverged": int(np.count_nonzero(info))}


def cpu_baseline(a_host, seconds=10.0):
    """The reference's CPU path restated in C (oracle/): EISPACK tred1 + imtql2
    + trbak on every thread of the host's share (one matrix per OpenMP thread),
    the same on one thread, and the classical Jacobi of jac.cpp:60-217 on every
    thread; each leg bounded to about `seconds` of CPU work."""
    import oracle
    threads, visible = host_threads()
    L = a_host.shape[0]
    allc = time_cpu(a_host, oracle.IMTQL2, threads, seconds)
    one = time_cpu(a_host[: max(1, min(L, 512))], oracle.IMTQL2, 1, seconds / 2)
    jac = time_cpu(a_host[: max(1, min(L, 4096))], oracle.JACOBI_CLASSICAL, threads, seconds / 2)
    return {"value": allc["value"], "unit": "matrices/s", "cores": threads, "kind": "port",
            "host_cpus_visible": visible,
            "sample": f"{a_host.shape[0]} matrices of the same workload x {allc['passes']} passes "
                      f"({allc['seconds']:.1f} s), C restatement of EISPACK tred1 + imtql2 + trbak "
                      f"(oracle/), one matrix per OpenMP thread, {threads} threads",
            "nonconverged": allc["nonconverged"],
            "single_core": {"value": one["value"], "unit": "matrices/s", "cores": 1,
                            "sample": f"{one['matrices']} matrices ({one['seconds']:.1f} s)"},
            "jacobi_classical": {"value": jac["value"], "unit": "matrices/s", "cores": threads,
                                 "sample": f"{jac['matrices']} matrices ({jac['seconds']:.1f} s), "
                                           "classical max-pivot Jacobi of jac.cpp:60-217 restated in C",
                                 "nonconverged": jac["nonconverged"]}}


def accuracy_all(a0, z, w):
    """The reference checker's metrics (eigen_GPU_check.cu:34-81) for every
    matrix of the batch, on the GPU (eigen_GPU_check_x_*)."""
    import eigeng_batched as eg
    out = eg.check(a0, w, z)
    wmax = w.abs().amax(dim=1).double()
    return dict(matrices=int(out.shape[0]), checker_failures=int((out[:, 3] != 0).sum().item()),
                max_err1_rel=float((out[:, 0] / wmax).max().item()),
                max_orth=float(out[:, 1].max().item()), TOL=float(out[0, 2].item()))


def accuracy_sample(a0_host, z_host, w_host, count=256):
    import numpy as np
    import oracle
    L, n, nm = a0_host.shape
    idx = np.linspace(0, L - 1, min(count, L)).astype(int)
    worst = np.zeros(4)
    fails = 0
    for k in idx:
        a = a0_host[k].copy()
        out = oracle.check(a, w_host[k], z_host[k])
        fails += int(out[3] != 0)
        worst = np.maximum(worst, [out[0] / max(abs(w_host[k][0]), abs(w_host[k][-1])),
                                   out[1], out[4], out[5]])
    return dict(sample=len(idx), checker_failures=fails, max_err1_rel=float(worst[0]),
                max_orth=float(worst[1]), max_pair_residual=float(worst[2]),
                max_rel_residual_F=float(worst[3]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=16384)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--dtype", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--solver", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-n256", action="store_true", help="skip the n=256 FP64 secondary line")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch secondary line")
    ap.add_argument("--no-letkf48", action="store_true", help="skip the L=65536 n=48 line")
    ap.add_argument("--no-n128", action="store_true", help="skip the n=128 FP32 line")
    ap.add_argument("--n128-only", action="store_true", help="print only the n=128 FP32 line")
    ap.add_argument("--no-n32", action="store_true", help="skip the n=32 FP64 L=1024 line")
    ap.add_argument("--n32-only", action="store_true", help="print only the n=32 FP64 L=1024 line")
    ap.add_argument("--no-letkf-step", action="store_true", help="skip the LETKF analysis-step line")
    ap.add_argument("--no-vendor", action="store_true", help="skip the hipSOLVER syevjBatched comparison")
    ap.add_argument("--no-tuned", action="store_true", help="skip the autotuned lines")
    ap.add_argument("--letkf-step-only", action="store_true", help="print only the LETKF analysis-step line")
    ap.add_argument("--n256-only", action="store_true",
                    help="run only the n=256 FP64 line (for profiling its kernels)")
    ap.add_argument("--no-warmstart", action="store_true", help="skip the learned warm-start line")
    ap.add_argument("--warmstart-only", action="store_true", help="print only the learned warm-start line")
    ap.add_argument("--ws-train-steps", type=int, default=50, help="warm-start line: training steps")
    ap.add_argument("--ws-total-steps", type=int, default=200, help="warm-start line: stream length")
    ap.add_argument("--ws-eval-steps", type=int, nargs="+", default=[50, 100, 150, 199],
                    help="warm-start line: evaluation steps")
    ap.add_argument("--no-chunks", action="store_true", help="skip the chunk trade-off line")
    ap.add_argument("--no-mixed", action="store_true",
                    help="n256 line without the mixed-precision and chunk sub-lines (for profiling)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="print only the n=256 FP64 mixed-precision line (and the plain time)")
    ap.add_argument("--uneven-count", type=int, default=100000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: matrices are independent, "
                                   "a multi-GPU run is L/N contiguous slices with no exchange; "
                                   "--gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    import eigeng_batched as eg

    if args.warmstart_only:
        print(json.dumps({"warmstart": bench_warmstart(train_steps=args.ws_train_steps,
                                                       total_steps=args.ws_total_steps,
                                                       eval_steps=tuple(args.ws_eval_steps))}))
        return
    if args.mixed_only:
        import torch
        q0 = make_qdq_gpu(8192, 256, seed=0xE16E0004, dtype=torch.float64)
        plain_ms, _, a2, w2, _ = time_solver(q0, 0, args.steps, args.warmup)
        del a2, w2
        print(json.dumps({"mixed": bench_mixed(q0, plain_ms, args.steps, args.warmup), "plain_ms": plain_ms}))
        return
    if args.n256_only:
        print(json.dumps({"n256": bench_n256(args.steps, args.warmup, vendor=not args.no_vendor,
                                             tuned=not args.no_tuned, mixed=not args.no_mixed)}))
        return
    if args.letkf_step_only:
        print(json.dumps({"letkf_step": bench_letkf_step(args.steps, args.warmup)}))
        return
    if args.n128_only:
        print(json.dumps({"n128": bench_n128(args.steps, args.warmup)}))
        return
    if args.n32_only:
        print(json.dumps({"n32": bench_n32(args.steps, args.warmup, tuned=not args.no_tuned)}))
        return
    dt = torch.float32 if args.dtype == "fp32" else torch.float64
    L, n = args.L, args.n
    k_mem, p_obs, rho, sigma = n, 400, 1.1, 0.1
    a0 = make_letkf_gpu(L, k_mem, p_obs, rho, sigma, seed=0xE16E0002, dtype=dt)
    ms, wall, a, w, info = time_solver(a0, args.solver, args.steps, args.warmup)
    mats = L / (ms * 1e-3)

    nbad = int((info != 0).sum().item())
    phases = None
    if n <= 64 and L <= 16384 and L >= 2048 and args.solver != 2:
        phases, _, _, _ = time_phases(a0, args.steps, 1)
    z_host = a.cpu().numpy()
    w_host = w.cpu().numpy()
    a0_host = a0.cpu().numpy()
    acc = accuracy_sample(a0_host, z_host, w_host)

    s = 4 if args.dtype == "fp32" else 8
    flops = (10.0 / 3.0) * n ** 3
    bytes_alg = s * (n * (n + 1) / 2 + n * n + n)
    peak_c = PEAK_FP32_TFLOPS if args.dtype == "fp32" else PEAK_FP64_TFLOPS
    if phases:
        # the step is a sequence of kernels: the roof is taken over their summed
        # event-timed durations (the fallback launch included in "apply")
        kern_ms = sum(phases.values())
        kernel = "eig_tri_kernel + eig_ql_reg_kernel + eig_apply_kernel (+ overflow-fallback launch)"
    else:
        kern_ms = ms
        kernel = "one solver kernel per step"
    achieved_tf = L * flops / (kern_ms * 1e-3) / 1e12
    achieved_gbs = L * bytes_alg / (kern_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_eig_{'split' if phases else 'wave'}_n{n}_{args.dtype}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    out = {
        "metric": "matrices/s at n=64,256 (tuned vs plain, equal residual); % HBM/MFMA peak",
        "value": mats,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.dtype == "fp32" else "f64",
        "data": "synthetic (seeded GPU generator, LETKF ensemble-space matrices)",
        "config": {"workload": f"plain solver (Householder + IMTQL2 + back-transform), L={L} "
                               f"LETKF matrices n=nm=m={n} {args.dtype}, k={k_mem} p={p_obs} "
                               f"rho={rho} sigma={sigma}",
                   "L": L, "n": n, "solver": ["imtql2", "tql2", "jacobi"][args.solver]},
        "roofline": {"bound": "valu", "achieved": achieved_tf, "peak": peak_c, "unit": "TFLOP/s",
                     "frac": achieved_tf / peak_c, "traffic": traffic, "kernel": kernel,
                     "kernel_ms": kern_ms, "phase_ms": phases,
                     "note": "compute roof at the reference's (10/3)n^3 flops per matrix over the "
                             "event-timed kernel time; the kernels run on the VALU (no MFMA), so "
                             "the peak is the FP vector peak"},
        "roofline_hbm": {"achieved": achieved_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved_gbs / PEAK_HBM_GBS, "bytes_per_matrix": bytes_alg},
        "wall_s_including_restores": wall,
        "nonconverged": nbad,
        "accuracy": acc,
    }
    out["accuracy_all"] = accuracy_all(a0, a, w)
    if n <= 64 and L >= 2048 and not args.no_chunks:
        out["chunk_tradeoff"] = chunk_tradeoff(a0, (0, 1, 2, 3), steps=args.steps, warmup=args.warmup)
    if not args.no_tuned:
        out["tuned"] = tuned_line(a0, ms, args.steps, args.warmup)
    if not args.no_vendor:
        out["vendor_baseline"] = vendor_baseline(a0)
    if not args.no_n256:
        out["n256"] = bench_n256(3, 1, vendor=not args.no_vendor, tuned=not args.no_tuned)
    if not args.no_n128:
        out["n128"] = bench_n128(5, 1)
    if not args.no_n32:
        out["n32"] = bench_n32(20, 3, tuned=not args.no_tuned)
    if not args.no_letkf48:
        out["letkf48"] = bench_letkf48(5, 1)
    if not args.no_letkf_step:
        out["letkf_step"] = bench_letkf_step(10, 2)
    if not args.no_warmstart:
        out["warmstart"] = bench_warmstart()
    if not args.no_uneven:
        del a, a0
        torch.cuda.empty_cache()
        out["uneven"] = bench_uneven(args.uneven_count, 8, 128, 3, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a0_host[: min(L, 16384)], args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
