"""The plain FP64 GPU VMC step: the denominator of the speed-up target.

This is the reference algorithm written directly in PyTorch FP64 on the GPU,
with no kernel of netket_amd:

* Metropolis sampling with a full forward pass of the RBM per proposal
  (netket/sampler/metropolis.py:416-464), vectorised over chains and
  sequential over proposals; exchange proposals pick the k-th differing
  cluster uniformly and carry the log n(x) - log n(x') correction
  (netket/sampler/rules/exchange.py:143-184), local proposals flip a uniform
  site (netket/sampler/rules/local.py:40-49);
* local energies from the materialised connected configurations and a
  batched log psi over all of them (netket/vqs/mc/kernels.py:62-71), zero
  matrix elements included as the padded get_conn does;
* the dense Jacobian O (Ns x Np, FP64), centred and scaled by 1/sqrt(Ns)
  (netket/jax/_jacobian/logic.py:308-387);
* S = O_L^T O_L + eps I and F = O_L^T dv as FP64 GEMMs (rocBLAS), the
  Cholesky factorisation and solve (rocSOLVER), and the SGD update
  (netket/_src/ngd/sr.py:16-80, netket/optimizer/solver/solvers.py:241-270).

The proposal loop (hundreds of small launches per step) is captured once in
a HIP graph and replayed, as a jit-compiled scan would run it; the rest of
the step runs eagerly.  Only bench.py uses this file, to time the baseline.
"""

from __future__ import annotations

import math

import numpy as np
import torch

LN2 = math.log(2.0)


def log_cosh(x):
    ax = x.abs()
    return ax + torch.log1p(torch.exp(-2.0 * ax)) - LN2


class PlainRBMVMC:
    """SR-VMC of an RBM (alpha * N hidden units) on a spin-1/2 chain, FP64.

    ``ham`` is ("heisenberg", J) (Marshall sign rule, Pauli matrices, the
    nearest-neighbour bonds of the periodic chain, Sz = 0) or
    ("ising", J, h) (periodic chain)."""

    def __init__(self, N, alpha, n_chains, n_samples, n_discard, ham, diag_shift, lr, seed=0,
                 init_std=0.01, sweep_size=None, device="cuda"):
        self.N, self.M = N, alpha * N
        self.C = n_chains
        self.Ns = n_samples
        self.Lc = n_samples // n_chains
        self.D = n_discard
        self.S = N if sweep_size is None else sweep_size
        self.ham = ham
        self.eps = diag_shift
        self.lr = lr
        self.dev = torch.device(device)
        g = np.random.default_rng(seed)
        M = self.M
        self.b = torch.tensor(g.normal(scale=init_std, size=M), dtype=torch.float64, device=self.dev)
        self.W = torch.tensor(g.normal(scale=init_std, size=(N, M)), dtype=torch.float64, device=self.dev)
        self.a = torch.tensor(g.normal(scale=init_std, size=N), dtype=torch.float64, device=self.dev)
        self.bonds = torch.tensor([[i, (i + 1) % N] for i in range(N)], dtype=torch.long, device=self.dev)
        if ham[0] == "heisenberg":
            sig = np.ones((n_chains, N))
            for c in range(n_chains):
                sig[c, g.permutation(N)[: N // 2]] = -1
        else:
            sig = g.choice([-1.0, 1.0], size=(n_chains, N))
        self.sigma = torch.tensor(sig, dtype=torch.float64, device=self.dev)
        self.samples = torch.empty((n_chains, self.Lc, N), dtype=torch.float64, device=self.dev)
        self.ar = torch.arange(n_chains, device=self.dev)
        self.lp = torch.empty(n_chains, dtype=torch.float64, device=self.dev)
        self._graph = None

    # -- model ------------------------------------------------------------------
    def log_psi(self, x):
        return log_cosh(x @ self.W + self.b).sum(-1) + x @ self.a

    # -- sampler ----------------------------------------------------------------
    def _proposal(self):
        x = self.sigma
        u = torch.rand(self.C, 2, dtype=torch.float64, device=self.dev)
        xp = x.clone()
        if self.ham[0] == "heisenberg":
            i, j = self.bonds[:, 0], self.bonds[:, 1]
            mask = x[:, i] != x[:, j]
            n = mask.sum(1)
            k = torch.clamp((u[:, 0] * n).floor().long(), max=self.N - 1)
            pos = (mask.long().cumsum(1) > k[:, None]).long().argmax(1)
            bi, bj = i[pos], j[pos]
            xp[self.ar, bi] = x[self.ar, bj]
            xp[self.ar, bj] = x[self.ar, bi]
            n_new = (xp[:, i] != xp[:, j]).sum(1)
            corr = torch.log(n.double()) - torch.log(n_new.double())
        else:
            site = torch.clamp((u[:, 0] * self.N).floor().long(), max=self.N - 1)
            xp[self.ar, site] = -x[self.ar, site]
            corr = 0.0
        lpp = 2.0 * self.log_psi(xp)
        acc = u[:, 1] < torch.exp(lpp - self.lp + corr)
        self.sigma.copy_(torch.where(acc[:, None], xp, x))
        self.lp.copy_(torch.where(acc, lpp, self.lp))

    def _sample_body(self):
        self.lp.copy_(2.0 * self.log_psi(self.sigma))
        for sw in range(self.D + self.Lc):
            for _ in range(self.S):
                self._proposal()
            if sw >= self.D:
                self.samples[:, sw - self.D].copy_(self.sigma)

    def sample(self, use_graph=True):
        if not use_graph:
            self._sample_body()
            return self.samples
        if self._graph is None:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self._sample_body()  # warm-up outside the capture
            torch.cuda.current_stream().wait_stream(s)
            self._graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self._graph):
                self._sample_body()
        self._graph.replay()
        return self.samples

    # -- local energies ---------------------------------------------------------
    def local_energy(self, x):
        """E_loc from the materialised connected configurations (x: (Ns, N))."""
        Ns, N = x.shape
        i, j = self.bonds[:, 0], self.bonds[:, 1]
        zz = (x[:, i] * x[:, j]).sum(1)
        lp = self.log_psi(x)
        if self.ham[0] == "heisenberg":
            J = self.ham[1]
            nb = i.numel()
            xp = x[:, None, :].repeat(1, nb, 1)
            b = torch.arange(nb, device=self.dev)
            xp[:, b, i] = x[:, j]
            xp[:, b, j] = x[:, i]
            mel = torch.where(x[:, i] != x[:, j], -2.0 * J, 0.0)  # sign rule: -2J on antiparallel bonds
            diag = J * zz
        else:
            J, h = self.ham[1], self.ham[2]
            xp = x[:, None, :].repeat(1, N, 1)
            s = torch.arange(N, device=self.dev)
            xp[:, s, s] = -x[:, s]
            mel = torch.full((Ns, N), -h, dtype=torch.float64, device=self.dev)
            diag = J * zz  # H = J sum z_i z_j - h sum x_i (netket/operator/_ising/jax.py:125-165)
        lpp = self.log_psi(xp.reshape(-1, N)).reshape(Ns, -1)
        return diag + (mel * torch.exp(lpp - lp[:, None])).sum(1)

    # -- the step ---------------------------------------------------------------
    def jacobian(self, x):
        t = torch.tanh(x @ self.W + self.b)
        Ns = x.shape[0]
        return torch.cat([t, (x[:, :, None] * t[:, None, :]).reshape(Ns, -1), x], dim=1)

    def step(self, use_graph=True):
        x = self.sample(use_graph).reshape(-1, self.N)
        e = self.local_energy(x)
        Ns = x.shape[0]
        O = self.jacobian(x)
        OL = (O - O.mean(0)) / math.sqrt(Ns)
        del O
        emean = e.mean()
        dv = 2.0 * (e - emean) / math.sqrt(Ns)
        Smat = OL.T @ OL
        Smat.diagonal().add_(self.eps)
        F = OL.T @ dv
        L, info = torch.linalg.cholesky_ex(Smat)
        dp = torch.cholesky_solve(F[:, None], L)[:, 0]
        M, N = self.M, self.N
        self.b -= self.lr * dp[:M]
        self.W -= self.lr * dp[M: M + N * M].reshape(N, M)
        self.a -= self.lr * dp[M + N * M:]
        return emean
