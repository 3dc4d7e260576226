"""CPU restatement of NetKet's VMC step — oracle / bench baseline (test infrastructure only).

One VMC_SR step in FP64 NumPy + LAPACK following the reference algorithm
(not the GPU's): a full RBM forward per Metropolis proposal
(netket/sampler/metropolis.py:427-460), materialised connected elements and
a batched log psi for E_loc (netket/vqs/mc/kernels.py:62-71), the dense
per-sample Jacobian (netket/jax/_jacobian/logic.py:356-373), S = O_L^T O_L
(netket/_src/ngd/sr.py:43-63) and cho_factor/cho_solve
(netket/optimizer/solver/solvers.py:241-270), then the SGD update.
"""

from __future__ import annotations

import time

import numpy as np
import scipy.linalg as sla

from . import operators as O
from . import stats as ostats
from .rbm import RBMParams, jacobian, log_psi


def exchange_sweeps(p: RBMParams, sigma: np.ndarray, clusters: np.ndarray, n_sweeps: int,
                    sweep_size: int, rng: np.random.Generator, record_from: int):
    """Vectorised over chains: exchange proposals with the n_diff correction."""
    C, N = sigma.shape
    x = sigma.copy()
    lp = 2.0 * log_psi(p, x)
    out = []
    rows = np.arange(C)
    for sw in range(n_sweeps):
        for _ in range(sweep_size):
            mask = x[:, clusters[:, 0]] != x[:, clusters[:, 1]]
            nd = mask.sum(axis=1)
            k = np.floor(rng.random(C) * nd).astype(np.int64)
            csum = np.cumsum(mask, axis=1)
            sel = np.argmax(csum > k[:, None], axis=1)
            u, v = clusters[sel, 0], clusters[sel, 1]
            xp = x.copy()
            xp[rows, u], xp[rows, v] = x[rows, v], x[rows, u]
            nd2 = (xp[:, clusters[:, 0]] != xp[:, clusters[:, 1]]).sum(axis=1)
            lpp = 2.0 * log_psi(p, xp)
            acc = rng.random(C) < np.exp(lpp - lp + np.log(nd) - np.log(nd2))
            x = np.where(acc[:, None], xp, x)
            lp = np.where(acc, lpp, lp)
        if sw >= record_from:
            out.append(x.copy())
    return np.stack(out, axis=1), x


def sr_stages(flat: np.ndarray, N: int, M: int, x: np.ndarray, terms: dict, diag_shift: float,
              timings: dict | None = None):
    """E_loc, S, F and the SR update on given samples x (Ns, N), FP64.
    Returns (e, S, F, dp)."""
    t1 = time.perf_counter()
    p = RBMParams(flat, N, M)
    xp, mels, _ = O.local_conn(x, terms)
    B, K, _ = xp.shape
    lpx = log_psi(p, x)
    lpp = log_psi(p, xp.reshape(-1, N)).reshape(B, K)
    e = np.sum(mels * np.exp(lpp - lpx[:, None]), axis=1)
    t2 = time.perf_counter()
    Ns = x.shape[0]
    Ojac = jacobian(p, x)
    OL = (Ojac - Ojac.mean(axis=0, keepdims=True)) / np.sqrt(Ns)
    del Ojac
    dv = 2.0 * (e - e.mean()) / np.sqrt(Ns)
    F = OL.T @ dv
    S = OL.T @ OL
    del OL
    S[np.diag_indices_from(S)] += diag_shift
    t3 = time.perf_counter()
    dp = sla.cho_solve(sla.cho_factor(S, lower=False), F)
    t4 = time.perf_counter()
    if timings is not None:
        for k, v in (("local_energy", t2 - t1), ("jacobian_gram", t3 - t2), ("cholesky", t4 - t3)):
            timings[k] = timings.get(k, 0.0) + v
    return e, S, F, dp


def vmc_sr_step(flat: np.ndarray, N: int, M: int, sigma: np.ndarray, terms: dict,
                clusters: np.ndarray, *, n_discard: int, chain_length: int, sweep_size: int,
                diag_shift: float, lr: float, rng: np.random.Generator, timings: dict | None = None):
    """One SR step; returns (new flat params, new chain state, energy mean)."""
    t0 = time.perf_counter()
    p = RBMParams(flat, N, M)
    samples, sigma = exchange_sweeps(p, sigma, clusters, n_discard + chain_length, sweep_size, rng,
                                     n_discard)
    t1 = time.perf_counter()
    if timings is not None:
        timings["sample"] = timings.get("sample", 0.0) + (t1 - t0)
    C, L, _ = samples.shape
    x = samples.reshape(-1, N)
    e, S, F, dp = sr_stages(flat, N, M, x, terms, diag_shift, timings)
    st = ostats.statistics(e.reshape(C, L))
    return flat - lr * dp, sigma, st["mean"]
