"""Graphs that feed the operators and the exchange rule of the VMC path.

Only the order of edges matters downstream: it fixes the order of the
Heisenberg connected elements (netket/operator/_graph_operator.py:131-143) and
the exchange cluster list (netket/sampler/rules/exchange.py:190-205).  The
reference builds the nearest-neighbour edges of a lattice in
``get_nn_edges`` (netket/graph/_lattice_edge_logic.py:102-137):

1. lattice sites are extended by ``order`` unit cells along periodic axes and
   enumerated cell-major (``np.mgrid`` order, netket/graph/_lattice_edge_logic.py:26-48);
2. pairs of extended sites ``a < b`` are grouped by their (binned) distance, the
   k-th group holding the k-th smallest distance, each group sorted
   lexicographically (netket/graph/_lattice_edge_logic.py:87-99);
3. each pair is mapped to real site ids, oriented ``(min, max)`` and inserted
   into a CPython ``set``; the edges of order k are the set's iteration order
   (netket/graph/_lattice_edge_logic.py:121-137).

This module does the same three steps on integer lattice coordinates (so the
distance binning is exact) and keeps step 3 as a real CPython ``set`` of int
tuples, whose iteration order depends only on the int hashes and the insertion
sequence.  The graph keeps edges in insertion order, oriented ``(min, max)``,
as the reference's undirected igraph edge list does (netket/graph/graph.py:151-175).
"""

from __future__ import annotations

from collections import deque
from typing import Sequence

import numpy as np


class Graph:
    """Undirected graph with coloured edges kept in insertion order."""

    def __init__(self, edges: Sequence, n_nodes: int | None = None):
        edges = list(edges)
        if len(edges) > 0 and len(edges[0]) == 3:
            colors = [int(c) for (_, _, c) in edges]
            edges = [(int(u), int(v)) for (u, v, _) in edges]
        else:
            colors = [0] * len(edges)
            edges = [(int(u), int(v)) for (u, v) in edges]
        if n_nodes is None:
            n_nodes = max((max(e) for e in edges), default=-1) + 1
        for u, v in edges:
            if not (0 <= u < n_nodes and 0 <= v < n_nodes):
                raise ValueError(f"edge {(u, v)} out of range for {n_nodes} nodes")
        # undirected edges are reported as (min, max)
        self._edges = [(min(u, v), max(u, v)) for (u, v) in edges]
        self._colors = colors
        self._n_nodes = int(n_nodes)

    @property
    def n_nodes(self) -> int:
        return self._n_nodes

    @property
    def n_edges(self) -> int:
        return len(self._edges)

    def nodes(self):
        return range(self._n_nodes)

    def edges(self, *, return_color: bool = False, filter_color: int | None = None):
        pairs = zip(self._edges, self._colors)
        if filter_color is not None:
            pairs = [(e, c) for (e, c) in pairs if c == filter_color]
        if return_color:
            return [(*e, c) for (e, c) in pairs]
        return [e for (e, _) in pairs]

    @property
    def edge_colors(self) -> list:
        return list(self._colors)

    def adjacency_list(self) -> list:
        adj = [[] for _ in range(self._n_nodes)]
        for u, v in self._edges:
            adj[u].append(v)
            if u != v:
                adj[v].append(u)
        return adj

    def distances(self) -> np.ndarray:
        """All-pairs hop distances by BFS (unreachable pairs are ``inf``)."""
        adj = self.adjacency_list()
        n = self._n_nodes
        d = np.full((n, n), np.inf)
        for s in range(n):
            d[s, s] = 0
            q = deque([s])
            while q:
                u = q.popleft()
                for w in adj[u]:
                    if d[s, w] == np.inf:
                        d[s, w] = d[s, u] + 1
                        q.append(w)
        return d

    def is_bipartite(self) -> bool:
        adj = self.adjacency_list()
        color = [-1] * self._n_nodes
        for s in range(self._n_nodes):
            if color[s] >= 0:
                continue
            color[s] = 0
            q = deque([s])
            while q:
                u = q.popleft()
                for w in adj[u]:
                    if w == u:
                        return False
                    if color[w] < 0:
                        color[w] = 1 - color[u]
                        q.append(w)
                    elif color[w] == color[u]:
                        return False
        return True

    def __repr__(self):
        return f"{type(self).__name__}(n_nodes={self.n_nodes}, n_edges={self.n_edges})"


def _lattice_edges(extent, pbc, order: int) -> list:
    """Coloured nearest-neighbour edges of a hypercubic Bravais lattice.

    Follows the three steps listed in the module docstring; the extended-site
    enumeration is row-major over the cell coordinates, and site ids are
    ``sum_d x_d * prod(extent[d+1:])``.
    """
    extent = np.asarray(extent, dtype=np.int64)
    ndim = extent.size
    pbc = np.broadcast_to(np.asarray(pbc, dtype=bool), (ndim,))
    shell = np.where(pbc, order, 0)
    axes = [np.arange(-shell[d], extent[d] + shell[d]) for d in range(ndim)]
    coords = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, ndim)
    radix = np.ones(ndim, dtype=np.int64)
    for d in range(ndim - 2, -1, -1):
        radix[d] = radix[d + 1] * extent[d + 1]
    ids = (coords % extent) @ radix

    # squared distances of all extended pairs a < b, up to (order)^2 (unit spacing)
    a_idx, b_idx = np.triu_indices(coords.shape[0], k=1)
    diff = coords[a_idx] - coords[b_idx]
    d2 = (diff * diff).sum(axis=1)
    keep = d2 <= order * order
    a_idx, b_idx, d2 = a_idx[keep], b_idx[keep], d2[keep]
    shells = np.unique(d2)

    colored = []
    for k in range(order):
        if k >= shells.size:
            break
        sel = d2 == shells[k]
        pairs = sorted(zip(a_idx[sel].tolist(), b_idx[sel].tolist()))
        true_edges = set()
        for n1, n2 in pairs:
            n1 = int(ids[n1])
            n2 = int(ids[n2])
            if n1 == n2:
                raise RuntimeError(
                    f"Lattice contains self-referential edge {(n1, n2)} of order {k}"
                )
            if n1 > n2:
                n1, n2 = n2, n1
            true_edges.add((n1, n2))
        for e in true_edges:
            colored.append((*e, k))
    return colored


class Lattice(Graph):
    """Hypercubic Bravais lattice (unit basis, one site per cell)."""

    def __init__(self, extent, *, pbc=True, max_neighbor_order: int | None = None):
        extent = [int(x) for x in extent]
        if any(x < 1 for x in extent):
            raise ValueError("extent must be positive")
        order = 1 if max_neighbor_order is None else int(max_neighbor_order)
        self._extent = np.asarray(extent)
        self._pbc = np.broadcast_to(np.asarray(pbc, dtype=bool), (len(extent),)).copy()
        self._max_neighbor_order = order
        colored = _lattice_edges(extent, self._pbc, order)
        super().__init__(colored, int(np.prod(extent)))

    @property
    def extent(self):
        return self._extent

    @property
    def pbc(self):
        return self._pbc

    @property
    def ndim(self) -> int:
        return self._extent.size

    @property
    def positions(self) -> np.ndarray:
        axes = [np.arange(e) for e in self._extent]
        return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, self.ndim).astype(float)


def Grid(extent, *, pbc=True, **kwargs) -> Lattice:
    return Lattice(extent, pbc=pbc, **kwargs)


def Hypercube(length: int, n_dim: int = 1, *, pbc: bool = True, **kwargs) -> Lattice:
    if not isinstance(length, (int, np.integer)) or length <= 0:
        raise TypeError("Argument `length` must be a positive integer")
    return Lattice([int(length)] * n_dim, pbc=pbc, **kwargs)


def Chain(length: int, *, pbc: bool = True, **kwargs) -> Lattice:
    return Hypercube(length, n_dim=1, pbc=pbc, **kwargs)


def Square(length: int, *, pbc: bool = True, **kwargs) -> Lattice:
    return Hypercube(length, n_dim=2, pbc=pbc, **kwargs)


def Cube(length: int, *, pbc: bool = True, **kwargs) -> Lattice:
    return Hypercube(length, n_dim=3, pbc=pbc, **kwargs)
