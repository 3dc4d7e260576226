"""Monte Carlo statistics of local estimators.

Follows the default (blocking) estimator ``mc_stats_old._statistics``
(netket/stats/mc_stats_old.py:87-196, selected by netket/stats/mc_stats.py:295-300).
The moments (mean, variance, variances of chain, block and half-chain means)
are reduced on the GPU in FP64 by ``nka_stats_moments``; the branch logic on
those eight numbers runs on the host when a field is first read, so building a
``Stats`` object does not synchronise the stream.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _native as nat
from . import _ops


class Stats:
    """mean, error_of_mean, variance, tau_corr, R_hat (netket/stats/mc_stats.py:84-180)."""

    def __init__(self, mean=np.nan, error_of_mean=np.nan, variance=np.nan, tau_corr=np.nan,
                 R_hat=np.nan, *, _moments=None, _shape=None):
        self._moments = _moments
        self._shape = _shape
        self._vals = None if _moments is not None else (mean, error_of_mean, variance, tau_corr, R_hat)
        self._host = None
        self._event = None

    def _resolve(self):
        if self._vals is None:
            if self._host is None:
                self._host = self._moments.cpu()
            m = self._host.numpy()
            self._vals = _finish(m, *self._shape)
        return self._vals

    @property
    def mean(self):
        return self._resolve()[0]

    @property
    def error_of_mean(self):
        return self._resolve()[1]

    @property
    def variance(self):
        return self._resolve()[2]

    @property
    def tau_corr(self):
        return self._resolve()[3]

    @property
    def R_hat(self):
        return self._resolve()[4]

    @property
    def mean_device(self) -> torch.Tensor:
        """The mean as a one-element FP64 device tensor (no synchronisation)."""
        return self._moments[0:1]

    def to_dict(self) -> dict:
        return {
            "Mean": self.mean,
            "Variance": self.variance,
            "Sigma": self.error_of_mean,
            "R_hat": self.R_hat,
            "TauCorr": self.tau_corr,
        }

    def to_compound(self):
        return "Mean", self.to_dict()

    def __repr__(self):
        m, e, v, t, r = self._resolve()
        return f"{m:.6f} ± {e:.6f} [σ²={v:.6f}, R̂={r:.4f}, τ={t:.1f}]"


def _finish(m: np.ndarray, C: int, L: int, batch_size: int = 32):
    """Branch logic of mc_stats_old._statistics on the reduced moments."""
    mean, var, cvar, n_blocks, bvar, hvar, l_block, ts = (float(x) for x in m[:8])
    n_blocks = int(n_blocks)
    l_block = int(l_block)
    n_batches = C
    with np.errstate(divide="ignore", invalid="ignore"):
        bare = np.float64(var)
        tau_batch = ((ts / n_batches) * np.float64(cvar) / bare - 1) * 0.5
        tau_block = ((ts / n_blocks) * np.float64(bvar) / bare - 1) * 0.5 if n_blocks > 0 else np.nan
    batch_good = bool(tau_batch < 6 * L) and n_batches >= batch_size
    block_good = bool(tau_block < 6 * l_block) and n_blocks >= batch_size
    if batch_good:
        err = math.sqrt(cvar / n_batches)
        tau = max(float(tau_batch), 0.0)
    elif block_good:
        err = math.sqrt(bvar / n_blocks)
        tau = max(float(tau_block), 0.0)
    else:
        err = float("nan")
        tau = float("nan")
    if n_batches > 1 and L >= 2:
        with np.errstate(divide="ignore", invalid="ignore"):
            r_hat = float(np.sqrt((L - 1) / L + np.float64(hvar) / bare))
    else:
        r_hat = float("nan")
    return mean, err, var, tau, r_hat


def statistics(data) -> Stats:
    """Statistics of ``data`` (n_chains, chain_length) or a 1-D series."""
    if not isinstance(data, torch.Tensor):
        data = torch.as_tensor(np.asarray(data))
    if data.ndim == 1:
        data = data.reshape(1, -1)
    if data.ndim != 2:
        raise NotImplementedError("Statistics are implemented only for ndim<=2")
    if data.is_complex():
        raise NotImplementedError("complex local estimators are not supported on the GPU path")
    d = data.to(device=_ops.device(), dtype=torch.float64).contiguous()
    C, L = d.shape
    out = torch.empty(8, dtype=torch.float64, device=d.device)
    with _ops.TIMER.region("stats"):
        nat.call("nka_stats_moments", _ops.ptr(d), C, L, _ops.ptr(out), _ops.stream())
    return Stats(_moments=out, _shape=(C, L))
