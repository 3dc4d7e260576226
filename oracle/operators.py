"""Connected elements and Hamiltonian matrices — oracle (test infrastructure only).

* ``ising_conn``  restates IsingJax.get_conn_padded
  (netket/operator/_ising/jax.py:125-165).
* ``local_conn``  restates LocalOperatorJax._get_conn_padded
  (netket/operator/_local_operator/jax.py:74-201) on a dict
  ``{acting_on (sorted tuple): matrix}`` in insertion order, with the packing
  rules of netket/operator/_local_operator/compile_helpers.py:29-218.
* ``dense_from_terms`` builds H independently by Kronecker products, the
  check that pins both.
* ``sparse_from_conn`` / ``lanczos`` restate nk.exact.lanczos_ed
  (netket/exact.py:24-97) on the oracle's own connected elements.

Basis convention: spin +1 <-> local index 0, spin -1 <-> local index 1; the
state index of a configuration has site 0 as the most significant bit.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

MEL_CUTOFF = 1e-10


# ---------------------------------------------------------------------------
# Ising
# ---------------------------------------------------------------------------
def ising_conn(x: np.ndarray, edges: np.ndarray, h: float, J: float, dtype=np.float64):
    """(xp, mels) for configurations x (B, N) of +-1 int8."""
    x = np.asarray(x)
    B, N = x.shape
    dtype = np.dtype(dtype)
    Jt, ht = dtype.type(J), dtype.type(h)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    same = x[:, edges[:, 0]] == x[:, edges[:, 1]]
    zz = (2 * same.astype(np.int64) - 1).sum(axis=1)
    if ht == 0:
        return x[:, None, :].copy(), (Jt * zz.astype(dtype)).reshape(B, 1).astype(dtype)
    K = N + 1
    xp = np.repeat(x[:, None, :], K, axis=1)
    for r in range(1, K):
        xp[:, r, r - 1] = -xp[:, r, r - 1]
    mels = np.empty((B, K), dtype=dtype)
    mels[:, 0] = Jt * zz.astype(dtype)
    mels[:, 1:] = -ht
    return xp, mels


# ---------------------------------------------------------------------------
# Local operators
# ---------------------------------------------------------------------------
def _row_index(xs_local: np.ndarray) -> np.ndarray:
    """Row of the term matrix from local indices of its sites (first site most significant)."""
    n = xs_local.shape[-1]
    w = 2 ** np.arange(n - 1, -1, -1)
    return (xs_local * w).sum(axis=-1)


def local_conn(x: np.ndarray, terms: dict, constant: float = 0.0, dtype=np.float64,
               mel_cutoff: float = MEL_CUTOFF, chunk: int = 32768):
    """(xp, mels, n_conn) for configurations x (B, N) and terms {sites: matrix}.

    Entry order before compaction: [diagonal if the operator has one], then for
    the groups of terms of equal acting size (ascending) and each term in
    insertion order, the off-diagonal elements of the selected row with
    |mel| > cutoff in column order, each term padded to its group's largest
    count with zero elements.  Kept are the entries with |mel| > cutoff, moved
    to the front in order; the tail is padded with (x, 0) up to max_conn_size.
    """
    x = np.asarray(x)
    B, N = x.shape
    dtype = np.dtype(dtype)
    items = [(tuple(s), np.asarray(m, dtype=dtype)) for s, m in terms.items()]
    sizes = sorted(set(len(s) for s, _ in items))
    nonzero_diag = abs(constant) >= mel_cutoff
    max_conn = 0
    tables = []  # per term: (sites, diag (R,), mel (R, nc), newloc (R, nc, g))
    for g in sizes:
        group = [(s, m) for s, m in items if len(s) == g]
        ncs = []
        for s, m in group:
            mask = np.abs(m) >= mel_cutoff
            np.fill_diagonal(mask, False)
            max_conn += int(mask.sum(axis=1).max())
            if np.any(np.abs(np.diag(m)) >= mel_cutoff):
                nonzero_diag = True
            ncs.append(max(int((np.abs(m - np.diag(np.diag(m))) > mel_cutoff).sum(axis=1).max()), 0))
        ncmax = max(ncs) if ncs else 0
        for s, m in group:
            R = m.shape[0]
            mel_t = np.zeros((R, ncmax), dtype=dtype)
            new_t = np.zeros((R, ncmax, g), dtype=np.int64)
            for r in range(R):
                k = 0
                for c in range(R):
                    if c != r and abs(m[r, c]) > mel_cutoff:
                        mel_t[r, k] = m[r, c]
                        new_t[r, k] = [(c >> (g - 1 - q)) & 1 for q in range(g)]
                        k += 1
            tables.append((g, np.array(s), np.diag(m).copy(), mel_t, new_t))
    if nonzero_diag:
        max_conn += 1
    K = max(max_conn, 1)
    xp_out = np.empty((B, K, N), dtype=x.dtype)
    mels_out = np.empty((B, K), dtype=dtype)
    n_out = np.empty(B, dtype=np.int32)
    for b0 in range(0, B, chunk):
        xb = x[b0 : b0 + chunk]
        nb = xb.shape[0]
        loc = (xb < 0).astype(np.int64)
        ent_m, ent_x = [], []
        if nonzero_diag:
            d = np.zeros(nb, dtype=dtype)
            for g in sizes:
                gs = np.zeros(nb, dtype=dtype)
                for (tg, s, dg, _, _) in tables:
                    if tg == g:
                        gs = gs + dg[_row_index(loc[:, s])]
                d = d + gs
            ent_m.append((dtype.type(constant) + d)[:, None])
            ent_x.append(xb[:, None, :])
        for (g, s, dg, mel_t, new_t) in tables:
            if mel_t.shape[1] == 0:
                continue
            r = _row_index(loc[:, s])
            ent_m.append(mel_t[r])
            xx = np.repeat(xb[:, None, :], mel_t.shape[1], axis=1)
            newvals = (1 - 2 * new_t[r]).astype(x.dtype)  # (nb, nc, g)
            valid = np.arange(mel_t.shape[1])[None, :] < (np.abs(mel_t[r]) > 0).sum(axis=1)[:, None]
            for q, site in enumerate(s):
                xx[:, :, site] = np.where(valid, newvals[:, :, q], xx[:, :, site])
            ent_x.append(xx)
        ent_m.append(np.zeros((nb, 1), dtype=dtype))
        ent_x.append(xb[:, None, :])
        mels = np.concatenate(ent_m, axis=1)
        xps = np.concatenate(ent_x, axis=1)
        keep = np.abs(mels) > mel_cutoff
        # indices of kept entries in order, padded with the final (x, 0) entry
        order = np.argsort(~keep, axis=1, kind="stable")[:, :K]
        kept_n = keep.sum(axis=1)
        pad = np.arange(K)[None, :] >= kept_n[:, None]
        order = np.where(pad, mels.shape[1] - 1, order)
        xp_out[b0 : b0 + nb] = np.take_along_axis(xps, order[:, :, None], axis=1)
        mels_out[b0 : b0 + nb] = np.take_along_axis(mels, order, axis=1)
        n_out[b0 : b0 + nb] = kept_n
    return xp_out, mels_out, n_out


def canonical_terms(operators, acting_on) -> dict:
    """{sorted sites: matrix} with duplicate site sets summed, insertion order kept
    (netket/operator/_local_operator/helpers.py:75-133, base.py:136-148)."""
    out = {}
    for m, s in zip(operators, acting_on):
        m = np.asarray(m, dtype=np.float64)
        s = tuple(int(v) for v in s)
        if list(s) != sorted(s):
            m = embed(m, s, tuple(sorted(s)))
            s = tuple(sorted(s))
        out[s] = out[s] + m if s in out else m
    return out


def embed(mat: np.ndarray, sites: tuple, union: tuple) -> np.ndarray:
    """Matrix acting on ``sites`` expressed on the ordered site tuple ``union``."""
    n_u = len(union)
    pos = [union.index(s) for s in sites]
    out = np.zeros((2**n_u, 2**n_u), dtype=mat.dtype)
    for r in range(2**n_u):
        rb = [(r >> (n_u - 1 - k)) & 1 for k in range(n_u)]
        rl = 0
        for p in pos:
            rl = 2 * rl + rb[p]
        for cl in range(2 ** len(sites)):
            v = mat[rl, cl]
            if v == 0:
                continue
            cb = list(rb)
            for j, p in enumerate(pos):
                cb[p] = (cl >> (len(sites) - 1 - j)) & 1
            c = 0
            for k in range(n_u):
                c = 2 * c + cb[k]
            out[r, c] += v
    return out


# ---------------------------------------------------------------------------
# Hamiltonian matrices
# ---------------------------------------------------------------------------
SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
I2 = np.eye(2)


def kron_site(op: np.ndarray, sites, N: int) -> np.ndarray:
    """Dense 2^N matrix of a k-site operator by explicit Kronecker products."""
    sites = list(sites)
    k = len(sites)
    if k == 1:
        mats = [op if i == sites[0] else I2 for i in range(N)]
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(out, m)
        return out
    # general: expand op in the basis of products of 2x2 unit matrices
    out = np.zeros((2**N, 2**N))
    units = [np.array([[1.0, 0], [0, 0]]), np.array([[0, 1.0], [0, 0]]),
             np.array([[0, 0], [1.0, 0]]), np.array([[0, 0], [0, 1.0]])]
    for r in range(2**k):
        for c in range(2**k):
            v = op[r, c]
            if v == 0:
                continue
            mats = [I2] * N
            for j, s in enumerate(sites):
                rb = (r >> (k - 1 - j)) & 1
                cb = (c >> (k - 1 - j)) & 1
                mats = list(mats)
                mats[s] = units[2 * rb + cb]
            prod = mats[0]
            for m in mats[1:]:
                prod = np.kron(prod, m)
            out += v * prod
    return out


def dense_ising(N: int, edges, h: float, J: float) -> np.ndarray:
    H = np.zeros((2**N, 2**N))
    for i in range(N):
        H -= h * kron_site(SX, [i], N)
    for u, v in edges:
        H += J * kron_site(SZ, [u], N) @ kron_site(SZ, [v], N)
    return H


def dense_from_terms(N: int, operators, acting_on, constant: float = 0.0) -> np.ndarray:
    H = constant * np.eye(2**N)
    for m, s in zip(operators, acting_on):
        H += kron_site(np.asarray(m, dtype=np.float64), s, N)
    return H


def states_all(N: int, n_down: int | None = None) -> np.ndarray:
    idx = np.arange(2**N, dtype=np.int64)
    bits = (idx[:, None] >> np.arange(N - 1, -1, -1)[None, :]) & 1
    x = (1 - 2 * bits).astype(np.int8)
    if n_down is not None:
        x = x[bits.sum(axis=1) == n_down]
    return x


def state_index(x: np.ndarray) -> np.ndarray:
    N = x.shape[-1]
    bits = (np.asarray(x) < 0).astype(np.int64)
    return (bits * (2 ** np.arange(N - 1, -1, -1))).sum(axis=-1)


def sparse_from_conn(states: np.ndarray, conn_fn, chunk: int = 32768) -> sp.csr_matrix:
    """Sparse H on the given basis states from a get_conn_padded-like function."""
    n = states.shape[0]
    keys = state_index(states)
    order = np.argsort(keys)
    skeys = keys[order]
    rows, cols, vals = [], [], []
    for s0 in range(0, n, chunk):
        xs = states[s0 : s0 + chunk]
        xp, mels = conn_fn(xs)[:2]
        B, K, N = xp.shape
        k2 = state_index(xp.reshape(-1, N)).reshape(B, K)
        pos = np.searchsorted(skeys, k2)
        pos = np.minimum(pos, n - 1)
        found = (skeys[pos] == k2) & (np.asarray(mels) != 0)
        bi, ki = np.nonzero(found)
        rows.append(s0 + bi)
        cols.append(order[pos[bi, ki]])
        vals.append(np.asarray(mels, dtype=np.float64)[bi, ki])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def lanczos(H: sp.spmatrix, k: int = 1) -> np.ndarray:
    """Lowest eigenvalues (netket/exact.py:65-97 calls eigsh(which='SA'))."""
    w = spla.eigsh(H, k=k, which="SA", return_eigenvectors=False)
    return np.sort(w)


def tfim_chain_e0(L: int, h: float, J: float = 1.0) -> float:
    """Closed-form ground energy of the periodic TFIM chain."""
    n = np.arange(L)
    return float(-np.sum(np.sqrt(J * J + h * h - 2 * J * h * np.cos(np.pi * (2 * n + 1) / L))))
