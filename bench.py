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


def make_qdq_gpu(L, n, seed, dtype, frac=0.05, gap=1e-8):
    """A = Q diag(lam) Q^T, Q from the QR of a Gaussian matrix, lam ~ U[-1,1]
    with a fraction `frac` of the eigenvalues replaced by pairs split by `gap`."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.empty((L, n, n), dtype=torch.float64, device="cuda")
    npairs = max(1, int(round(frac * n / 2)))
    chunk = 64                      # batched QR workspace grows with the batch
    for s in range(0, L, chunk):
        e = min(L, s + chunk)
        G = torch.randn((e - s, n, n), generator=g, device="cuda", dtype=torch.float64)
        Q, _ = torch.linalg.qr(G)
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


def cpu_baseline(a_host, min_seconds=10.0, max_reps=20):
    import numpy as np
    import oracle
    threads = os.cpu_count() or 1
    threads = min(threads, 16)
    reps, t0, total = 0, time.perf_counter(), 0
    while True:
        _, _, info = oracle.eig_batch(a_host, oracle.IMTQL2, threads=threads)
        reps += 1
        total += a_host.shape[0]
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            break
    return total / el, threads, reps, el, int(np.count_nonzero(info))


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: matrices are independent, "
                                   "a multi-GPU run is L/N contiguous slices with no exchange "
                                   "(see DESIGN.md); --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    import eigeng_batched as eg

    dt = torch.float32 if args.dtype == "fp32" else torch.float64
    L, n = args.L, args.n
    k_mem, p_obs, rho, sigma = n, 400, 1.1, 0.1
    a0 = make_letkf_gpu(L, k_mem, p_obs, rho, sigma, seed=0xE16E0002, dtype=dt)
    ms, wall, a, w, info = time_solver(a0, args.solver, args.steps, args.warmup)
    mats = L / (ms * 1e-3)

    nbad = int((info != 0).sum().item())
    z_host = a.cpu().numpy()
    w_host = w.cpu().numpy()
    a0_host = a0.cpu().numpy()
    acc = accuracy_sample(a0_host, z_host, w_host)

    s = 4 if args.dtype == "fp32" else 8
    flops = (10.0 / 3.0) * n ** 3
    bytes_alg = s * (n * (n + 1) / 2 + n * n + n)
    peak_c = PEAK_FP32_TFLOPS if args.dtype == "fp32" else PEAK_FP64_TFLOPS
    achieved_tf = mats * flops / 1e12
    achieved_gbs = mats * bytes_alg / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_eig_wave_n{n}_{args.dtype}.json")
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
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": peak_c, "unit": "TFLOP/s",
                     "frac": achieved_tf / peak_c, "traffic": traffic,
                     "note": "compute roof at the reference's (10/3)n^3 flops per matrix; the "
                             "kernel runs on the VALU (no MFMA), so this is the FP vector peak"},
        "roofline_hbm": {"achieved": achieved_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved_gbs / PEAK_HBM_GBS, "bytes_per_matrix": bytes_alg},
        "wall_s_including_restores": wall,
        "nonconverged": nbad,
        "accuracy": acc,
    }
    if not args.no_n256:
        L2, n2 = 8192, 256
        q0 = make_qdq_gpu(L2, n2, seed=0xE16E0004, dtype=torch.float64)
        ms2, _, a2, w2, info2 = time_solver(q0, args.solver, 2, 1)
        acc2 = accuracy_sample(q0.cpu().numpy(), a2.cpu().numpy(), w2.cpu().numpy(), count=32)
        del q0, a2
        out["n256"] = {"workload": "L=8192 n=256 FP64, A = Q diag(lam) Q^T, lam ~ U[-1,1], 5% pairs "
                                   "split by 1e-8", "value": L2 / (ms2 * 1e-3), "unit": "matrices/s",
                       "ms_per_step": ms2, "nonconverged": int((info2 != 0).sum().item()),
                       "accuracy": acc2,
                       "roofline_fp64": {"achieved": L2 / (ms2 * 1e-3) * (10.0 / 3.0) * n2 ** 3 / 1e12,
                                         "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}}
    if not args.no_cpu:
        sample = a0_host[: min(L, 16384)]
        v, cores, reps, el, cbad = cpu_baseline(sample, args.cpu_seconds)
        out["cpu_baseline"] = {"value": v, "unit": "matrices/s", "cores": cores, "kind": "port",
                               "sample": f"{sample.shape[0]} matrices of the same workload x {reps} "
                                         f"passes ({el:.1f} s), C oracle (EISPACK tred1+imtql2+"
                                         f"trbak restatement), one matrix per OpenMP thread"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
