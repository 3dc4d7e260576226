"""Two-layer dense network ansatz — oracle (test infrastructure only).

The network is the user flax module of the J1-J2 workload (the pattern of
docs/tutorials/gs-heisenberg.ipynb "Model2" with two Dense layers):
``Dense(H1) -> log_cosh -> Dense(H2) -> log_cosh -> sum`` with ``nn.Dense``
semantics (netket/nn/blocks/mlp.py:105-115: x @ kernel + bias) and log_cosh
as netket/nn/activation.py:78-84.  Parameters in the ravel order of the
flax dict (keys sorted): Dense_0/bias (H1), Dense_0/kernel (N, H1),
Dense_1/bias (H2), Dense_1/kernel (H1, H2).

* ``log_psi``   the forward pass in FP64.
* ``jacobian``  per-sample log-derivatives by the chain rule (pinned in the
                tests against central finite differences of ``log_psi``).
* ``local_energy`` the reference algorithm of local_value_kernel_jax
                (netket/vqs/mc/kernels.py:62-71): a full forward of every
                connected configuration.
"""

from __future__ import annotations

import numpy as np

from .rbm import log_cosh


class FFNNParams:
    """Unpacked parameters (FP64) in the flat ravel order."""

    def __init__(self, flat: np.ndarray, N: int, H1: int, H2: int, use_bias: bool = True):
        flat = np.asarray(flat, dtype=np.float64)
        self.N, self.H1, self.H2 = N, H1, H2
        off = 0
        self.b1 = None
        if use_bias:
            self.b1 = flat[off : off + H1]
            off += H1
        self.W1 = flat[off : off + N * H1].reshape(N, H1)
        off += N * H1
        self.b2 = None
        if use_bias:
            self.b2 = flat[off : off + H2]
            off += H2
        self.W2 = flat[off : off + H1 * H2].reshape(H1, H2)
        off += H1 * H2
        self.n_params = off


def forward(p: FFNNParams, x):
    """(theta1, h1, z2) of a batch of configurations."""
    x = np.asarray(x, dtype=np.float64)
    th1 = x @ p.W1
    if p.b1 is not None:
        th1 = th1 + p.b1
    h1 = log_cosh(th1)
    z2 = h1 @ p.W2
    if p.b2 is not None:
        z2 = z2 + p.b2
    return th1, h1, z2


def log_psi(p: FFNNParams, x):
    return log_cosh(forward(p, x)[2]).sum(axis=-1)


def jacobian(p: FFNNParams, x) -> np.ndarray:
    """O[s, k] = d log psi(x_s) / d theta_k in ravel order, FP64 (Ns, Np)."""
    x = np.asarray(x, dtype=np.float64)
    th1, h1, z2 = forward(p, x)
    g2 = np.tanh(z2)                    # d logpsi / d z2
    d1 = (g2 @ p.W2.T) * np.tanh(th1)   # d logpsi / d theta1
    ns = x.shape[0]
    blocks = []
    if p.b1 is not None:
        blocks.append(d1)
    blocks.append((x[:, :, None] * d1[:, None, :]).reshape(ns, -1))
    if p.b2 is not None:
        blocks.append(g2)
    blocks.append((h1[:, :, None] * g2[:, None, :]).reshape(ns, -1))
    return np.concatenate(blocks, axis=1)


def jacobian_fd(p: FFNNParams, flat: np.ndarray, x, eps: float = 1e-6, use_bias: bool = True):
    """Central finite differences of log_psi with respect to every parameter."""
    flat = np.asarray(flat, dtype=np.float64)
    out = np.empty((np.asarray(x).shape[0], flat.size))
    for k in range(flat.size):
        fp = flat.copy()
        fp[k] += eps
        fm = flat.copy()
        fm[k] -= eps
        lp = log_psi(FFNNParams(fp, p.N, p.H1, p.H2, use_bias), x)
        lm = log_psi(FFNNParams(fm, p.N, p.H1, p.H2, use_bias), x)
        out[:, k] = (lp - lm) / (2 * eps)
    return out


def local_energy(p: FFNNParams, x, conn_fn):
    """E_loc(s) = sum_k mel_k exp(log psi(x'_k) - log psi(s)) (reference algorithm)."""
    x = np.asarray(x)
    xp, mels = conn_fn(x)[:2]
    B, K, N = xp.shape
    lp = log_psi(p, x)
    lpp = log_psi(p, xp.reshape(-1, N)).reshape(B, K)
    return np.sum(np.asarray(mels, dtype=np.float64) * np.exp(lpp - lp[:, None]), axis=-1)
