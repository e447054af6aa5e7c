"""GPU: the warm-start network's kernels against the float64 CPU reference
(oracle/warmnet_ref.py), and the learned warm start end to end.

Bars: forward output and loss within 1e-4 relative, the gradient within
1e-3 ||g|| (the layers multiply on the matrix cores with FP32 operands, the
reference in float64); training lowers the loss; the ablation reports what the
trained network changes against the identity start (Jacobi sweeps, fallbacks)
while every warm-started result passes the reference checker
(eigen_GPU_check.cu:34-81).
"""
import numpy as np
import pytest
import torch

import eigeng_batched as eg
import oracle
import warmnet_ref
import warmstart
import workloads as wl

pytestmark = pytest.mark.gpu
DEV = "cuda"


def batch_and_targets(B, n, seed):
    a0, _ = wl.letkf_y(B, k=n, p=60, rho=1.05, dtype=np.float32, seed=seed)
    a = torch.from_numpy(a0).to(DEV)
    v, w, info = eg.eigen_batch(a.clone())
    torch.cuda.synchronize()
    return a, v


def margins(p, A, V, n, h):
    """Distance of the ReLU inputs from 0 and of the sign-fixing choice from a
    tie, relative to their scale: the gradient is not continuous there, so a
    float32 and a float64 evaluation may legitimately differ."""
    W1, b1, W2, b2, W3, b3 = warmnet_ref.unpack(p, n, h)
    x = warmnet_ref.features(A)
    z1 = x @ W1 + b1
    z2 = np.maximum(z1, 0) @ W2 + b2
    m_relu = min(np.abs(z1).min() / np.abs(z1).max(), np.abs(z2).min() / np.abs(z2).max())
    s = np.sort(np.abs(V), axis=1)
    m_tie = float(np.min(s[:, -1, :] - s[:, -2, :]))
    return m_relu, m_tie


@pytest.mark.parametrize("n,h,B", [(3, 8, 5), (6, 32, 16), (48, 64, 70)])
def test_gradient_matches_the_cpu_reference(n, h, B):
    """Parameters drawn on the host (the first seed whose ReLU inputs and sign
    choices keep a margin from their kinks: 1e-4 relative for the ReLU inputs, no
    exact tie for the signs), then the GPU's loss,
    gradient and forward output against the float64 reference."""
    print(f"[warmnet] gradient n={n} h={h} B={B}", flush=True)
    net = warmstart.WarmNet(n, hidden=h, batch=B, seed=n)
    a, v = batch_and_targets(B, n, seed=n + 1)
    A = np.swapaxes(a.cpu().numpy().astype(np.float64), 1, 2)      # row-major matrices
    V = np.swapaxes(v.cpu().numpy().astype(np.float64), 1, 2)      # column j = eigenvector j
    _, _, offs = warmnet_ref.dims(n, h)
    for seed in range(100):
        p = (np.random.default_rng(seed).normal(size=offs[-1]) * 0.3).astype(np.float32).astype(np.float64)
        mr, mt = margins(p, A, V, n, h)
        if mr > 1e-4 and mt > 0.0:
            break
    else:
        pytest.skip("no parameter draw with margins from the kinks")
    net.p.copy_(torch.from_numpy(p.astype(np.float32)))
    loss = float(net.grad(a, v).item())
    g = net.g.cpu().numpy().astype(np.float64)
    lr, gr = warmnet_ref.loss_and_grad(p, A, V, n, h)
    assert abs(loss - lr) <= 1e-4 * max(lr, 1e-30), (loss, lr)
    assert np.linalg.norm(g - gr) <= 1e-3 * np.linalg.norm(gr), np.linalg.norm(g - gr) / np.linalg.norm(gr)
    q0 = net.forward(a)
    torch.cuda.synchronize()
    _, _, _, y = warmnet_ref.forward(p, A, n, h)
    qy = q0.cpu().numpy()[:, :, :n].reshape(B, n * n)             # column-major n x n per matrix
    assert np.max(np.abs(qy - y)) <= 1e-4 * max(np.abs(y).max(), 1.0)


def test_init_is_he_normal():
    """eigen_GPU_warmnet_init: weights with std sqrt(2 / fan_in), zero biases."""
    n, h = 16, 256
    net = warmstart.WarmNet(n, hidden=h, batch=4, seed=11)
    p = net.p.cpu().numpy().astype(np.float64)
    W1, b1, W2, b2, W3, b3 = warmnet_ref.unpack(p, n, h)
    din = n * (n + 1) // 2
    for W, fan in ((W1, din), (W2, h), (W3, h)):
        assert abs(W.std() / np.sqrt(2.0 / fan) - 1.0) < 0.05
        assert abs(W.mean()) < 0.05 * np.sqrt(2.0 / fan)
    assert not b1.any() and not b2.any() and not b3.any()


def test_training_lowers_the_loss():
    print("[warmnet] training", flush=True)
    n, h, B = 16, 64, 256
    net = warmstart.WarmNet(n, hidden=h, batch=B, seed=3)
    a, v = batch_and_targets(B, n, seed=9)
    first = float(net.train_step(a, v, lr=1e-3).item())
    for _ in range(200):
        last = float(net.train_step(a, v, lr=1e-3).item())
    assert last < 0.5 * first, (first, last)


def test_learned_warm_start_ablation():
    """Train on the plain solver's eigenpairs of the first steps of a drifting
    stream, then solve later steps from the network's Q0, from the identity and
    from the previous step's eigenvectors.  Every result passes the checker; the
    sweep counts and fallbacks of each start are reported (what the network
    gains, or not, against the method without it)."""
    print("[warmnet] ablation", flush=True)
    n, L = 24, 4096
    stream = warmstart.DriftingStream(L, n, p=100, seed=5)
    net = warmstart.WarmNet(n, hidden=128, batch=1024, seed=1)
    hist = warmstart.train_on_stream(net, stream, steps=6, epochs=4, lr=1e-3)
    assert hist[-1] < hist[0]
    a = stream.matrices()
    prev = a.clone()
    prev, _, _ = eg.eigen_batch(prev)          # the previous step's eigenvectors (temporal start)
    stream.advance()
    a = stream.matrices()
    zp, wp, _ = eg.eigen_batch(a.clone())
    starts = {"network": net.forward(a),
              "identity": torch.eye(n, device=DEV).expand(L, n, n).contiguous(),
              "previous_step": prev}
    report = {}
    a_np = a.cpu().numpy()
    full = wl.unpack(a_np)
    for name, q0 in starts.items():
        z, w, info, sw = warmstart.solve_warm(a, q0, max_sweeps=10)
        torch.cuda.synchronize()
        assert int(info.abs().max()) == 0
        swn = sw.cpu().numpy()
        report[name] = (float(swn.mean()), int((swn >= 10).sum()))
        wn, zn = w.cpu().numpy(), z.cpu().numpy()
        assert np.max(np.abs(wn - wp.cpu().numpy()) / np.abs(wn).max(axis=1, keepdims=True)) <= 31 * 2 ** -23 * np.sqrt(n + 16)
        for k in range(0, L, 512):
            out = oracle.check(np.ascontiguousarray(full[k].T), wn[k], zn[k])
            assert int(out[3]) == 0, (name, k)
    print("[warmnet] mean sweeps / fallbacks:", report, flush=True)
    # the previous step's eigenvectors are a strong start on a slowly drifting stream
    assert report["previous_step"][0] < report["identity"][0]
