"""Operators of the VMC path: transverse-field Ising and local (bond) operators.

Host side of the operator plug-in API ``get_conn_padded(x) -> (xp, mels)`` and
``max_conn_size`` (netket/operator/_discrete_operator_jax.py:112-139).  The
connected elements are generated on the GPU by the HIP kernels
``nka_ising_conn`` and ``nka_localop_conn``; the same operator descriptors feed
the fused local-energy kernel, which never materialises ``xp``.

* ``Ising`` follows ``IsingJax`` (netket/operator/_ising/jax.py:36-165,
  netket/operator/_ising/base.py:32-161).
* ``LocalOperator`` follows ``LocalOperatorJax``: operators are canonicalised to
  sorted ``acting_on`` (netket/operator/_local_operator/helpers.py:75-133),
  terms acting on the same sites are summed (netket/operator/_local_operator/base.py:136-148),
  and the lookup tables are packed per group of equal acting size
  (netket/operator/_local_operator/compile_helpers.py:29-218).
* ``Heisenberg`` / ``GraphOperator`` build the bond operators as the reference does
  (netket/operator/_heisenberg.py:33-132, netket/operator/_graph_operator.py:47-145).
"""

from __future__ import annotations

import itertools
from numbers import Number

import numpy as np

from .hilbert import Spin

MEL_CUTOFF = 1.0e-10
MAX_ACTING_SIZE = 4


def _canon_dtype(*vals, dtype=None):
    if dtype is not None:
        dtype = np.dtype(dtype)
        if dtype.kind == "c":
            raise NotImplementedError("complex operators are not supported on the MI355X path")
        return dtype
    for v in vals:
        if np.iscomplexobj(v):
            raise NotImplementedError("complex operators are not supported on the MI355X path")
    return np.dtype(np.float64)


# ---------------------------------------------------------------------------
# Ising
# ---------------------------------------------------------------------------
class Ising:
    r"""Transverse-field Ising Hamiltonian :math:`-h\sum_i \sigma^x_i + J\sum_{<ij>}\sigma^z_i\sigma^z_j`.

    Connected elements follow ``_ising_kernel_jax`` (netket/operator/_ising/jax.py:147-165):
    row 0 is ``x`` itself with mel ``J * sum_edges(2*[x_u == x_v] - 1)``; row ``r``
    (1 <= r <= N) is ``x`` with site ``r-1`` flipped and mel ``-h``.  When ``h == 0``
    only row 0 is produced.
    """

    is_hermitian = True

    def __init__(self, hilbert: Spin, graph, h: float, J: float = 1.0, dtype=None):
        if hilbert.local_size != 2:
            raise ValueError("Ising only supports Spin-1/2 hilbert spaces.")
        dtype = _canon_dtype(h, J, dtype=dtype)
        if hasattr(graph, "edges") and callable(graph.edges):
            if graph.n_nodes != hilbert.size:
                raise ValueError("The size of the graph must match the hilbert space.")
            edges = graph.edges()
        else:
            edges = graph
        edges = np.asarray([[int(u), int(v)] for (u, v) in edges], dtype=np.int64).reshape(-1, 2)
        if edges.ndim != 2 or edges.shape[1] != 2:
            raise ValueError("Graph should be a graph, a list of edges or an (E,2) array")
        if edges.size and (edges.min() < 0 or edges.max() >= hilbert.size):
            raise ValueError("edge index out of range")
        self.hilbert = hilbert
        self._h = dtype.type(h)
        self._J = dtype.type(J)
        self._dtype = dtype
        self._edges = edges
        self._device_cache = {}

    @property
    def h(self):
        return self._h

    @property
    def J(self):
        return self._J

    @property
    def edges(self) -> np.ndarray:
        return self._edges

    @property
    def dtype(self):
        return self._dtype

    @property
    def max_conn_size(self) -> int:
        return self.hilbert.size + 1

    @property
    def _conn_rows(self) -> int:
        """Rows actually produced by get_conn_padded (1 when h == 0)."""
        return 1 if self._h == 0 else self.hilbert.size + 1

    def get_conn_padded(self, x):
        from . import _ops

        return _ops.ising_conn(self, x)

    def n_conn(self, x):
        from . import _ops

        return _ops.ising_n_conn(self, x)

    def to_local_operator(self) -> "LocalOperator":
        ha = LocalOperator(self.hilbert, dtype=self._dtype)
        if self._h != 0:
            for i in range(self.hilbert.size):
                ha -= float(self._h) * sigmax(self.hilbert, i, dtype=self._dtype)
        if self._J != 0:
            for i, j in self._edges:
                ha += float(self._J) * (
                    sigmaz(self.hilbert, int(i), dtype=self._dtype)
                    @ sigmaz(self.hilbert, int(j), dtype=self._dtype)
                )
        return ha

    def __repr__(self):
        return f"Ising(J={self._J}, h={self._h}; dim={self.hilbert.size})"


# ---------------------------------------------------------------------------
# Local operators
# ---------------------------------------------------------------------------
def _embed(mat: np.ndarray, sites: tuple, union: tuple) -> np.ndarray:
    """Matrix of ``mat`` (acting on ``sites``) on the sorted site tuple ``union``."""
    n_u = len(union)
    out = np.zeros((2**n_u, 2**n_u), dtype=mat.dtype)
    pos = [union.index(s) for s in sites]
    rest = [k for k in range(n_u) if k not in pos]
    for r in range(2**n_u):
        rb = [(r >> (n_u - 1 - k)) & 1 for k in range(n_u)]
        r_loc = 0
        for p in pos:
            r_loc = 2 * r_loc + rb[p]
        for c_loc in range(2 ** len(sites)):
            v = mat[r_loc, c_loc]
            if v == 0:
                continue
            cb = list(rb)
            for j, p in enumerate(pos):
                cb[p] = (c_loc >> (len(sites) - 1 - j)) & 1
            c = 0
            for k in range(n_u):
                c = 2 * c + cb[k]
            out[r, c] += v
    del rest
    return out


def _reorder(mat: np.ndarray, sites: tuple):
    """Permute ``mat`` so that its sites are sorted ascending (helpers.py:129)."""
    if list(sites) == sorted(sites):
        return mat, tuple(sites)
    srt = tuple(sorted(sites))
    return _embed(mat, tuple(sites), srt), srt


class LocalOperator:
    """Sum of few-site spin-1/2 operators given as dense matrices."""

    is_hermitian_cache = None

    def __init__(self, hilbert: Spin, operators=(), acting_on=(), constant=0.0, dtype=None,
                 *, mel_cutoff: float = MEL_CUTOFF):
        if isinstance(acting_on, Number):
            acting_on = [acting_on]
        acting_on = list(acting_on)
        operators = list(operators) if not isinstance(operators, np.ndarray) or operators.ndim == 3 else [operators]
        if acting_on and not any(hasattr(a, "__len__") for a in acting_on):
            operators = [np.asarray(operators[0]) if len(operators) == 1 else np.asarray(operators)]
            acting_on = [acting_on]
        ops = [np.asarray(o) for o in operators]
        self._dtype = _canon_dtype(*ops, constant, dtype=dtype)
        self.hilbert = hilbert
        self.mel_cutoff = float(mel_cutoff)
        self._constant = self._dtype.type(constant)
        self._ops = {}
        for op, aon in zip(ops, acting_on):
            aon = tuple(int(a) for a in aon)
            if len(aon) == 0:
                continue
            if max(aon) >= hilbert.size or min(aon) < 0:
                raise ValueError("An operator acts on an invalid set of sites.")
            if len(set(aon)) != len(aon):
                raise ValueError(f"The operator acts on duplicated sites {aon}")
            if op.shape != (2 ** len(aon), 2 ** len(aon)):
                raise ValueError(f"operator of shape {op.shape} does not match acting_on {aon}")
            if len(aon) > MAX_ACTING_SIZE:
                raise NotImplementedError(f"terms acting on more than {MAX_ACTING_SIZE} sites")
            op = op.astype(self._dtype)
            op, aon = _reorder(op, aon)
            self._add(aon, op)
        self._packed = None
        self._device_cache = {}

    def _add(self, aon, op):
        if aon in self._ops:
            op = self._ops[aon] + op
        self._ops[aon] = op

    # -- properties ----------------------------------------------------------
    @property
    def dtype(self):
        return self._dtype

    @property
    def constant(self):
        return self._constant

    @property
    def operators(self):
        return list(self._ops.values())

    @property
    def acting_on(self):
        return [list(a) for a in self._ops.keys()]

    @property
    def n_operators(self):
        return len(self._ops)

    @property
    def is_hermitian(self) -> bool:
        return all(np.allclose(o, o.T.conj()) for o in self._ops.values())

    @property
    def max_conn_size(self) -> int:
        return self.pack()["max_conn_size"]

    # -- algebra ---------------------------------------------------------------
    def copy(self, *, dtype=None):
        new = LocalOperator(self.hilbert, dtype=dtype or self._dtype, mel_cutoff=self.mel_cutoff)
        new._constant = new._dtype.type(self._constant)
        new._ops = {k: v.astype(new._dtype) for k, v in self._ops.items()}
        return new

    def __iadd__(self, other):
        if isinstance(other, Number):
            self._constant = self._dtype.type(self._constant + other)
        elif isinstance(other, LocalOperator):
            if other.hilbert != self.hilbert:
                raise ValueError("operators act on different hilbert spaces")
            self._constant = self._dtype.type(self._constant + other._constant)
            for aon, op in other._ops.items():
                self._add(aon, op.astype(self._dtype))
        else:
            return NotImplemented
        self._packed = None
        self._device_cache = {}
        return self

    def __add__(self, other):
        res = self.copy()
        res += other
        return res

    __radd__ = __add__

    def __neg__(self):
        return -1.0 * self

    def __isub__(self, other):
        return self.__iadd__(-other)

    def __sub__(self, other):
        return self + (-other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, scalar):
        if not isinstance(scalar, Number):
            return NotImplemented
        res = self.copy()
        res._constant = res._dtype.type(res._constant * scalar)
        res._ops = {k: v * res._dtype.type(scalar) for k, v in res._ops.items()}
        return res

    __rmul__ = __mul__

    def __matmul__(self, other):
        if not isinstance(other, LocalOperator):
            return NotImplemented
        res = LocalOperator(self.hilbert, dtype=np.promote_types(self._dtype, other._dtype))
        terms_a = list(self._ops.items()) + ([((), np.array([[self._constant]]))] if self._constant != 0 else [])
        terms_b = list(other._ops.items()) + ([((), np.array([[other._constant]]))] if other._constant != 0 else [])
        for (sa, ma), (sb, mb) in itertools.product(terms_a, terms_b):
            union = tuple(sorted(set(sa) | set(sb)))
            if len(union) == 0:
                res._constant = res._dtype.type(res._constant + ma[0, 0] * mb[0, 0])
                continue
            ea = _embed(ma, sa, union) if sa else ma[0, 0] * np.eye(2 ** len(union))
            eb = _embed(mb, sb, union) if sb else mb[0, 0] * np.eye(2 ** len(union))
            res._add(union, (ea @ eb).astype(res._dtype))
        return res

    # -- packing (compile_helpers.py:29-218) ------------------------------------
    def pack(self) -> dict:
        """Flatten the operator into the lookup tables consumed by the kernels.

        Terms are grouped by acting size in ascending order (np.unique order of
        compile_helpers.py:199), keeping insertion order within a group.  For each
        term and each row ``r`` (the local indices of the term's sites, first site
        most significant, netket/operator/_local_operator/jax.py:37-55) the tables
        hold the diagonal element, the off-diagonal elements with ``|mel| > cutoff``
        in column order and the new local indices of those columns.
        """
        if self._packed is not None:
            return self._packed
        items = list(self._ops.items())
        sizes = sorted(set(len(a) for a, _ in items))
        ordered = [(a, m) for s in sizes for (a, m) in items if len(a) == s]
        op_sites, op_nsites, op_row_off = [], [], []
        row_diag, row_nconn, row_conn_off = [], [], []
        conn_mel, conn_newidx = [], []
        max_conn = 0
        any_diag = abs(self._constant) >= self.mel_cutoff
        for aon, mat in ordered:
            n = len(aon)
            op_sites.append(list(aon) + [-1] * (MAX_ACTING_SIZE - n))
            op_nsites.append(n)
            op_row_off.append(len(row_diag))
            mask = np.abs(mat) >= self.mel_cutoff
            np.fill_diagonal(mask, False)
            max_conn += int(mask.sum(axis=1).max()) if mat.shape[0] > 0 else 0
            if np.any(np.abs(np.diag(mat)) >= self.mel_cutoff):
                any_diag = True
            for r in range(2**n):
                row_diag.append(mat[r, r])
                row_conn_off.append(len(conn_mel))
                cnt = 0
                for c in range(2**n):
                    if c != r and abs(mat[r, c]) > self.mel_cutoff:
                        conn_mel.append(mat[r, c])
                        conn_newidx.append(c)
                        cnt += 1
                row_nconn.append(cnt)
        if any_diag:
            max_conn += 1
        dt = self._dtype
        self._packed = {
            "n_ops": len(ordered),
            "op_sites": np.asarray(op_sites, dtype=np.int32).reshape(-1, MAX_ACTING_SIZE),
            "op_nsites": np.asarray(op_nsites, dtype=np.int32),
            "op_row_off": np.asarray(op_row_off, dtype=np.int32),
            "row_diag": np.asarray(row_diag, dtype=dt),
            "row_nconn": np.asarray(row_nconn, dtype=np.int32),
            "row_conn_off": np.asarray(row_conn_off, dtype=np.int32),
            "conn_mel": np.asarray(conn_mel, dtype=dt),
            "conn_newidx": np.asarray(conn_newidx, dtype=np.int32),
            "constant": float(self._constant),
            "nonzero_diagonal": bool(any_diag),
            "max_conn_size": int(max_conn),
            "mel_cutoff": self.mel_cutoff,
        }
        return self._packed

    def get_conn_padded(self, x):
        from . import _ops

        return _ops.localop_conn(self, x)

    def n_conn(self, x):
        from . import _ops

        return _ops.localop_conn(self, x, return_n_conn=True)[2]

    def __repr__(self):
        return (
            f"LocalOperator(dim={self.hilbert.size}, #acting_on={self.n_operators} locations, "
            f"constant={self._constant}, dtype={self._dtype})"
        )


# ---------------------------------------------------------------------------
# spin operators (netket/operator/spin.py), real ones only
# ---------------------------------------------------------------------------
_SX = np.array([[0.0, 1.0], [1.0, 0.0]])
_SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
_SP = np.array([[0.0, 1.0], [0.0, 0.0]])  # sigma^+ in the (+1, -1) local basis
_SM = _SP.T.copy()


def sigmax(hilbert, site: int, dtype=None) -> LocalOperator:
    return LocalOperator(hilbert, [_SX], [[site]], dtype=dtype)


def sigmaz(hilbert, site: int, dtype=None) -> LocalOperator:
    return LocalOperator(hilbert, [_SZ], [[site]], dtype=dtype)


def sigmap(hilbert, site: int, dtype=None) -> LocalOperator:
    return LocalOperator(hilbert, [_SP], [[site]], dtype=dtype)


def sigmam(hilbert, site: int, dtype=None) -> LocalOperator:
    return LocalOperator(hilbert, [_SM], [[site]], dtype=dtype)


class spin:  # namespace mirroring ``netket.operator.spin``
    sigmax = staticmethod(sigmax)
    sigmaz = staticmethod(sigmaz)
    sigmap = staticmethod(sigmap)
    sigmam = staticmethod(sigmam)


# ---------------------------------------------------------------------------
# Graph operators
# ---------------------------------------------------------------------------
def GraphOperator(hilbert, graph, site_ops=(), bond_ops=(), bond_ops_colors=(), dtype=None):
    """Sum of site and bond operators over a graph (netket/operator/_graph_operator.py:103-145)."""
    if len(bond_ops) == 0 and len(site_ops) == 0:
        raise ValueError("Must input at least site_ops or bond_ops.")
    operators, acting_on = [], []
    for i in range(graph.n_nodes):
        for op in site_ops:
            operators.append(np.asarray(op))
            acting_on.append([i])
    if len(bond_ops_colors) > 0:
        if len(bond_ops) != len(bond_ops_colors):
            raise ValueError("The sizes of bond_ops and bond_ops_colors do not match.")
        for u, v, color in graph.edges(return_color=True):
            for c, bond_color in enumerate(bond_ops_colors):
                if bond_color == color:
                    operators.append(np.asarray(bond_ops[c]))
                    acting_on.append([u, v])
    elif len(bond_ops) > 0:
        assert len(bond_ops) == 1
        for u, v in graph.edges():
            operators.append(np.asarray(bond_ops[0]))
            acting_on.append([u, v])
    return LocalOperator(hilbert, operators, acting_on, dtype=dtype)


_SZSZ = np.array([[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]])
_EXCH = np.array([[0, 0, 0, 0], [0, 0, 2, 0], [0, 2, 0, 0], [0, 0, 0, 0]])


def Heisenberg(hilbert, graph, J=1.0, sign_rule=None, dtype=None) -> LocalOperator:
    """Heisenberg model on a graph (netket/operator/_heisenberg.py:33-132)."""
    if isinstance(J, (list, tuple)):
        if len(J) != max(graph.edge_colors) + 1:
            raise ValueError("the number of couplings must match the number of edge colours")
        if sign_rule is None:
            sign_rule = [False] * len(J)
        else:
            if len(sign_rule) != len(J):
                raise ValueError("sign_rule and J must have the same length")
            from .graph import Graph

            for i in range(len(J)):
                if sign_rule[i] and not Graph(graph.edges(filter_color=i)).is_bipartite():
                    raise ValueError("sign_rule=True specified for a non-bipartite lattice")
        bond_ops = [J[i] * (_SZSZ - _EXCH if sign_rule[i] else _SZSZ + _EXCH) for i in range(len(J))]
        colors = list(range(len(J)))
    else:
        if sign_rule is None:
            sign_rule = graph.is_bipartite()
        elif sign_rule and not graph.is_bipartite():
            raise ValueError("sign_rule=True specified for a non-bipartite lattice")
        bond_ops = [J * (_SZSZ - _EXCH if sign_rule else _SZSZ + _EXCH)]
        colors = []
    return GraphOperator(hilbert, graph, bond_ops=bond_ops, bond_ops_colors=colors, dtype=dtype)
