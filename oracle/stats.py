"""Blocking statistics — oracle (test infrastructure only).

Restates ``mc_stats_old._statistics`` (netket/stats/mc_stats_old.py:87-196),
the estimator selected by default in netket/stats/mc_stats.py:295-300.
"""

from __future__ import annotations

import numpy as np


def statistics(data, batch_size: int = 32) -> dict:
    data = np.atleast_1d(np.asarray(data, dtype=np.float64))
    if data.ndim == 1:
        data = data.reshape(1, -1)
    C, L = data.shape
    mean = data.mean()
    variance = data.var()
    ts = data.size
    chain_means = data.mean(axis=1)
    batch_var, n_batches = chain_means.var(), chain_means.size
    l_block = max(1, L // batch_size)
    nbc = L // l_block
    blocks = data[:, : nbc * l_block].reshape(-1, l_block).mean(axis=1)
    if blocks.size > 0:
        block_var, n_blocks = blocks.var(), blocks.size
    else:
        block_var, n_blocks = np.nan, 0
    with np.errstate(divide="ignore", invalid="ignore"):
        tau_batch = ((ts / n_batches) * batch_var / variance - 1) * 0.5
        tau_block = ((ts / n_blocks) * block_var / variance - 1) * 0.5 if n_blocks else np.nan
    batch_good = (tau_batch < 6 * L) and (n_batches >= batch_size)
    block_good = (tau_block < 6 * l_block) and (n_blocks >= batch_size)
    if batch_good:
        err, tau = np.sqrt(batch_var / n_batches), max(tau_batch, 0.0)
    elif block_good:
        err, tau = np.sqrt(block_var / n_blocks), max(tau_block, 0.0)
    else:
        err, tau = np.nan, np.nan
    if n_batches > 1 and L >= 2:
        if L % 2 == 0:
            halves = data.reshape(2 * C, L // 2)
        else:
            halves = data[:, :-1].reshape(2 * C, L // 2)
        split_var = halves.mean(axis=1).var()
        with np.errstate(divide="ignore", invalid="ignore"):
            r_hat = np.sqrt((L - 1) / L + split_var / variance)
    else:
        r_hat = np.nan
    return {"mean": mean, "error_of_mean": err, "variance": variance, "tau_corr": tau,
            "R_hat": r_hat}
