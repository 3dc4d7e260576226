"""RBM ansatz, local energies and log-derivatives — oracle (test infrastructure only).

* ``log_psi``       RBM.__call__ (netket/models/rbm.py:57-81) with log_cosh
                    exactly as netket/nn/activation.py:78-84, in FP64.
* ``local_energy``  local_value_kernel_jax (netket/vqs/mc/kernels.py:62-71):
                    get_conn_padded, a full forward of every connected
                    configuration, sum mel * exp(dlogpsi).  This is the
                    reference algorithm, not the GPU's hidden-unit update.
* ``jacobian``      per-sample log-derivatives in the ravel order
                    Dense/bias, Dense/kernel, visible_bias
                    (netket/jax/_jacobian/logic.py:308-387), from the closed
                    form of the RBM derivative.
* ``forces_to_grad`` / ``expect_and_grad`` forces_expect_hermitian +
                    force_to_grad (netket/vqs/mc/mc_state/expect_forces.py:67-112,
                    netket/vqs/mc/common.py:102-118), one weighted formula
                    whose uniform-weight case is the sample estimate.
"""

from __future__ import annotations

import numpy as np

from . import stats as ostats


def log_cosh(x):
    x = np.asarray(x, dtype=np.float64)
    ax = np.abs(x)
    return ax + np.log1p(np.exp(-2.0 * ax)) - np.log(2.0)


class RBMParams:
    """Unpacked RBM parameters (FP64) with the flat ravel order."""

    def __init__(self, flat: np.ndarray, N: int, M: int, hidden_bias: bool = True,
                 visible_bias: bool = True):
        flat = np.asarray(flat, dtype=np.float64)
        self.N, self.M = N, M
        off = 0
        self.b = None
        if hidden_bias:
            self.b = flat[off : off + M]
            off += M
        self.W = flat[off : off + N * M].reshape(N, M)
        off += N * M
        self.a = None
        if visible_bias:
            self.a = flat[off : off + N]
            off += N
        self.n_params = off


def theta(p: RBMParams, x):
    x = np.asarray(x, dtype=np.float64)
    th = x @ p.W
    if p.b is not None:
        th = th + p.b
    return th


def log_psi(p: RBMParams, x):
    x = np.asarray(x, dtype=np.float64)
    out = log_cosh(theta(p, x)).sum(axis=-1)
    if p.a is not None:
        out = out + x @ p.a
    return out


def local_energy(p: RBMParams, x, conn_fn):
    """E_loc(s) = sum_k mel_k exp(log psi(x'_k) - log psi(s)) (reference algorithm)."""
    x = np.asarray(x)
    xp, mels = conn_fn(x)[:2]
    B, K, N = xp.shape
    lp = log_psi(p, x)
    lpp = log_psi(p, xp.reshape(-1, N)).reshape(B, K)
    return np.sum(np.asarray(mels, dtype=np.float64) * np.exp(lpp - lp[:, None]), axis=-1)


def jacobian(p: RBMParams, x) -> np.ndarray:
    """O[s, k] = d log psi(x_s) / d theta_k in ravel order, FP64 (Ns, Np)."""
    x = np.asarray(x, dtype=np.float64)
    t = np.tanh(theta(p, x))
    blocks = []
    if p.b is not None:
        blocks.append(t)
    blocks.append((x[:, :, None] * t[:, None, :]).reshape(x.shape[0], -1))
    if p.a is not None:
        blocks.append(x)
    return np.concatenate(blocks, axis=1)


def forces_to_grad(O, eloc, weights=None):
    """grad = 2 Re sum_s w_s O_s^T conj(E_s - <E>_w), <E>_w = sum_s w_s E_s.

    forces_expect_hermitian (netket/vqs/mc/mc_state/expect_forces.py:67-112)
    is the Monte Carlo estimate with w_s = 1/Ns over the samples, and
    force_to_grad (netket/vqs/mc/common.py:102-118) doubles the real part.
    The same formula with the exact weights w_s = |psi(s)|^2 / Z over the full
    basis is the gradient of the exact energy; the oracle tests pin it against
    central finite differences of <psi|H|psi>/<psi|psi>.  Without weights the
    uniform 1/Ns is used (the sample estimate)."""
    O = np.asarray(O, dtype=np.float64)
    eloc = np.asarray(eloc, dtype=np.float64)
    w = np.full(eloc.shape[0], 1.0 / eloc.shape[0]) if weights is None else np.asarray(weights, np.float64)
    ebar = np.sum(w * eloc)
    return 2.0 * (O.T @ (w * (eloc - ebar))), ebar


def expect_and_grad(p: RBMParams, samples, conn_fn):
    """(Stats, grad, E_loc) on the chain-major samples (C, L, N), the gradient
    by ``forces_to_grad`` with the uniform sample weights 1/Ns."""
    C, L, N = samples.shape
    x = samples.reshape(-1, N)
    e = local_energy(p, x, conn_fn)
    st = ostats.statistics(e.reshape(C, L))
    grad, _ = forces_to_grad(jacobian(p, x), e)
    return st, grad, e


def exact_probability(p: RBMParams, states) -> np.ndarray:
    """|psi|^2 / Z over the given basis states (FullSumState semantics,
    netket/vqs/full_summ/state.py:60-61, 336-365)."""
    lp = 2.0 * log_psi(p, states)
    lp -= lp.max()
    w = np.exp(lp)
    return w / w.sum()
