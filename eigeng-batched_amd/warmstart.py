"""Learned warm start for streams of LETKF-like matrices (n <= 64).

A 3-layer perceptron (eig_warmnet.hpp, behind eigen_GPU_warmnet_*) predicts an
approximate eigenbasis Q0 from a matrix; eigen_GPU_batch_warm_* makes it
orthonormal (batched Householder QR), rotates the matrix into it on the matrix
cores and finishes with a few Jacobi sweeps, with the plain solver (Householder
tridiagonalisation + QL) as the fallback whenever the sweep cap is reached.  The
network is trained on the GPU, by the library's own forward/backward kernels and
Adam, on eigenpairs the plain solver produces for the first steps of the stream.

Everything here drives the HIP library through its C ABI; torch only holds
device memory and streams.  The reference has no learned component: this is
a warm start for a stream of related batches, built around the reference's
Jacobi path (jac_tiled.hpp:3-152), which always starts from the identity.
"""
from __future__ import annotations

import ctypes

import eigeng_batched as eg


class WarmNet:
    """Parameters, Adam state and activation workspace of one network."""

    def __init__(self, n: int, hidden: int = 512, batch: int = 4096, seed: int = 0, device: str = "cuda"):
        import torch
        self.n, self.hidden, self.batch = n, hidden, batch
        lib = eg.lib()
        self.count = int(lib.eigen_GPU_warmnet_params(n, hidden))
        if self.count == 0:
            raise ValueError("n must be 1..64 and hidden 1..4096")
        f32 = dict(dtype=torch.float32, device=device)
        self.p = torch.empty(self.count, **f32)
        self.g = torch.zeros(self.count, **f32)
        self.m = torch.zeros(self.count, **f32)
        self.v = torch.zeros(self.count, **f32)
        self.loss = torch.zeros(1, **f32)
        self.ws = torch.empty(int(lib.eigen_GPU_warmnet_bufsize(batch, n, hidden)), dtype=torch.uint8, device=device)
        self.step = 0
        eg._check_rc(lib.eigen_GPU_warmnet_init(n, hidden, self.p.data_ptr(), ctypes.c_ulonglong(seed),
                                                eg._stream_ptr(None)), "eigen_GPU_warmnet_init")

    def forward(self, a, q0=None, stream=None):
        """q0 (a's layout) for the FP32 batch a (L, n, nm)."""
        import torch
        L, n, nm = a.shape
        if q0 is None:
            q0 = torch.empty_like(a)
        eg._check_rc(eg.lib().eigen_GPU_warmnet_forward(L, nm, n, self.hidden, self.p.data_ptr(), a.data_ptr(),
                                                        q0.data_ptr(), self.batch, self.ws.data_ptr(),
                                                        eg._stream_ptr(stream)), "eigen_GPU_warmnet_forward")
        return q0

    def grad(self, a, v, stream=None):
        """Loss (device tensor) and gradient (self.g) for a minibatch of at most `batch` matrices."""
        B, n, nm = a.shape
        if B > self.batch:
            raise ValueError("minibatch larger than the workspace's batch")
        eg._check_rc(eg.lib().eigen_GPU_warmnet_grad(B, nm, n, self.hidden, self.p.data_ptr(), a.data_ptr(),
                                                     v.data_ptr(), self.g.data_ptr(), self.loss.data_ptr(),
                                                     self.ws.data_ptr(), eg._stream_ptr(stream)),
                     "eigen_GPU_warmnet_grad")
        return self.loss

    def train_step(self, a, v, lr: float = 1e-3, stream=None):
        loss = self.grad(a, v, stream)
        self.step += 1
        eg._check_rc(eg.lib().eigen_GPU_warmnet_adam(self.count, self.p.data_ptr(), self.g.data_ptr(),
                                                     self.m.data_ptr(), self.v.data_ptr(), self.step, lr,
                                                     eg._stream_ptr(stream)), "eigen_GPU_warmnet_adam")
        return loss


class DriftingStream:
    """LETKF-like matrices A_t = ((k-1)/rho) I + Y_t^T Y_t, k = n ensemble
    members, p observations, Y_0 ~ N(0, sigma0^2), Y_t = Y_{t-1} + N(0, drift^2)
    (seeded, on the GPU; the products by eigen_GPU_batch_gram_FP on the matrix
    cores)."""

    def __init__(self, L: int, n: int, p: int = 300, rho: float = 1.05, sigma0: float = 0.1, drift: float = 0.01,
                 seed: int = 0xE16E0005, device: str = "cuda"):
        import torch
        self.L, self.n, self.p, self.rho, self.drift = L, n, p, rho, drift
        self.gen = torch.Generator(device=device)
        self.gen.manual_seed(seed)
        # y[k, i, o] = Y_k(o, i): the gram entry point's layout (ld p)
        self.y = torch.randn((L, n, p), generator=self.gen, device=device, dtype=torch.float32) * sigma0
        self.t = 0

    def matrices(self):
        """The current step's batch (L, n, n), FP32."""
        return eg.gram(self.y, self.p, alpha=(self.n - 1) / self.rho)

    def advance(self):
        import torch
        self.y += torch.randn(self.y.shape, generator=self.gen, device=self.y.device, dtype=self.y.dtype) * self.drift
        self.t += 1


def train_on_stream(net: WarmNet, stream: DriftingStream, steps: int, epochs: int = 1, lr: float = 1e-3,
                    log_every: int = 0):
    """Train on the plain solver's eigenpairs of `steps` consecutive steps of the
    stream (the stream advances).  Returns the loss history (one value per step,
    the mean over its minibatches)."""
    import torch
    hist = []
    for _ in range(steps):
        a = stream.matrices()
        v = a.clone()
        v, w, info = eg.eigen_batch(v)
        L = a.shape[0]
        tot = torch.zeros(1, device=a.device)
        nb = 0
        for _e in range(epochs):
            perm = torch.randperm(L, device=a.device)
            for s in range(0, L - net.batch + 1, net.batch):
                idx = perm[s:s + net.batch]
                tot += net.train_step(a[idx].contiguous(), v[idx].contiguous(), lr)
                nb += 1
        hist.append(float(tot.item()) / max(nb, 1))
        if log_every and len(hist) % log_every == 0:
            print(f"[warmstart] train step {len(hist)} loss {hist[-1]:.4e}", flush=True)
        stream.advance()
    return hist


def solve_warm(a, q0, max_sweeps: int = 6, workspace=None, stream=None):
    """Warm-started solve of a copy of a from q0; returns (z, w, info, sweeps)."""
    z = a.clone()
    return eg.eigen_batch_warm(z, q0, max_sweeps=max_sweeps, workspace=workspace, stream=stream)
