"""Metropolis sampling — oracle (test infrastructure only).

``metropolis`` restates MetropolisSampler._sample_chain/_sample_next
(netket/sampler/metropolis.py:416-505) with LocalRule
(netket/sampler/rules/local.py:40-49) or ExchangeRule
(netket/sampler/rules/exchange.py:143-218), driven by the same Philox4x32-10
draws as the GPU kernel (proposal q of chain c at counter (q, c, "METR")):
word 0 picks the site / the k-th differing cluster (with integer cluster
weights: the differing cluster whose cumulative weight first exceeds a draw
uniform in [0, total weight)), words 1-2 the uniform.
log psi is recomputed with a full forward per proposal (the reference
algorithm), in FP64.

``random_states`` restates the kernel's initial-state draw (uniform spins, or
a Fisher-Yates placement of the down spins).
"""

from __future__ import annotations

import numpy as np

from . import philox
from .rbm import log_psi


def random_states(C: int, N: int, n_down: int | None, seed: int, counter: int = 0) -> np.ndarray:
    out = np.empty((C, N), dtype=np.int8)
    for c in range(C):
        if n_down is None:
            x = np.empty(N, dtype=np.int8)
            for i0 in range(0, N, 128):
                r = philox.draw(counter + i0 // 128, c, philox.TAG_RANDSTATE, seed)
                for i in range(i0, min(N, i0 + 128)):
                    bit = (int(r[(i - i0) >> 5]) >> ((i - i0) & 31)) & 1
                    x[i] = -1 if bit else 1
        else:
            x = np.array([-1] * n_down + [1] * (N - n_down), dtype=np.int8)
            for i in range(N - 1, 0, -1):
                r = philox.draw(counter + i, c, philox.TAG_RANDSTATE, seed)
                j = philox.bounded(r[0], i + 1)
                x[i], x[j] = x[j], x[i]
        out[c] = x
    return out


def exchange_moves(x: np.ndarray, clusters: np.ndarray, weights=None):
    """Proposals of the exchange rule from x: (list of x', log-corrections,
    proposal probabilities).

    Each x' swaps the two sites of one cluster whose spins differ, drawn with
    probability w_c / W(x), W(x) the total weight of the differing clusters
    (w = 1 without weights); the correction is log W(x) - log W(x')
    (netket/sampler/rules/exchange.py:143-184, with ``probabilities`` as w).
    """
    w = np.ones(len(clusters)) if weights is None else np.asarray(weights, dtype=np.float64)
    diff = np.nonzero(x[clusters[:, 0]] != x[clusters[:, 1]])[0]
    tot = float(w[diff].sum())
    moves, corrs, probs = [], [], []
    for c in diff:
        u_, v_ = clusters[c]
        xp = x.copy()
        xp[u_], xp[v_] = x[v_], x[u_]
        tot_new = float(w[np.nonzero(xp[clusters[:, 0]] != xp[clusters[:, 1]])[0]].sum())
        moves.append(xp)
        corrs.append(np.log(tot) - np.log(tot_new))
        probs.append(w[c] / tot)
    return moves, corrs, probs


def pick_weighted(w: np.ndarray, k: int) -> int:
    """Index of the entry whose cumulative weight (in order) first exceeds k,
    0 <= k < sum(w); each index i is picked by exactly w[i] values of k."""
    return int(np.searchsorted(np.cumsum(np.asarray(w, dtype=np.int64)), k, side="right"))


def local_moves(x: np.ndarray):
    """Proposals of the local rule: flip one of the N sites, each with probability 1/N."""
    moves = []
    for i in range(x.size):
        xp = x.copy()
        xp[i] = -xp[i]
        moves.append(xp)
    return moves, [0.0] * x.size, [1.0 / x.size] * x.size


def metropolis(params, sigma0: np.ndarray, *, rule: str, clusters=None, n_discard: int,
               chain_length: int, sweep_size: int, seed: int, counter: int,
               machine_pow: float = 2.0, log_psi_fn=None, chain_ids=None, weights=None):
    """Returns (samples (C, L, N), final sigma, n_accepted per chain).  The
    ansatz is the RBM of ``params`` unless ``log_psi_fn(params, x)`` is given.
    ``chain_ids`` (default 0..C-1) are the chain indices of the rows of
    ``sigma0`` in the random streams, so a subset of a larger batch of chains
    can be followed."""
    log_psi_ = log_psi if log_psi_fn is None else log_psi_fn
    sigma = np.array(sigma0, dtype=np.int8, copy=True)
    C, N = sigma.shape
    samples = np.empty((C, chain_length, N), dtype=np.int8)
    nacc = np.zeros(C, dtype=np.int64)
    if clusters is not None:
        clusters = np.asarray(clusters, dtype=np.int64)
    ids = np.arange(C) if chain_ids is None else np.asarray(chain_ids)
    for row in range(C):
        c = int(ids[row])
        x = sigma[row].astype(np.int8).copy()
        lp = machine_pow * log_psi_(params, x[None])[0]
        q = counter
        for sw in range(n_discard + chain_length):
            for _ in range(sweep_size):
                r = philox.draw(q, c, philox.TAG_METROPOLIS, seed)
                q += 1
                if rule == "local":
                    k = philox.bounded(r[0], N)
                    xp = x.copy()
                    xp[k] = -xp[k]
                    corr = 0.0
                elif weights is None:
                    # the k-th differing cluster (ascending cluster order), as exchange_moves lists them
                    diff = np.nonzero(x[clusters[:, 0]] != x[clusters[:, 1]])[0]
                    if diff.size == 0:
                        continue
                    k = philox.bounded(r[0], diff.size)
                    u_, v_ = clusters[diff[k]]
                    xp = x.copy()
                    xp[u_], xp[v_] = x[v_], x[u_]
                    n_new = int(np.count_nonzero(xp[clusters[:, 0]] != xp[clusters[:, 1]]))
                    corr = np.log(diff.size) - np.log(n_new)
                else:
                    # integer weights: k uniform in [0, W(x)), the differing cluster
                    # whose cumulative weight (ascending cluster order) first exceeds k
                    wi = np.asarray(weights, dtype=np.int64)
                    diff = np.nonzero(x[clusters[:, 0]] != x[clusters[:, 1]])[0]
                    tot = int(wi[diff].sum())
                    if tot == 0:
                        continue
                    c_ = diff[pick_weighted(wi[diff], philox.bounded(r[0], tot))]
                    u_, v_ = clusters[c_]
                    xp = x.copy()
                    xp[u_], xp[v_] = x[v_], x[u_]
                    tot_new = int(wi[np.nonzero(xp[clusters[:, 0]] != xp[clusters[:, 1]])[0]].sum())
                    corr = np.log(float(tot)) - np.log(float(tot_new))
                lpp = machine_pow * log_psi_(params, xp[None])[0]
                u = philox.uniform01_f64(r)
                if u < np.exp(lpp - lp + corr):
                    x = xp
                    lp = lpp
                    nacc[row] += 1
            if sw >= n_discard:
                samples[row, sw - n_discard] = x
        sigma[row] = x
    return samples, sigma, nacc
