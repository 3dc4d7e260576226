"""Spin-1/2 Hilbert space of the sampled VMC path.

Mirrors ``netket.hilbert.Spin`` (netket/hilbert/spin.py:110-193) for s=1/2:

* local states are the spin values ``[+1, -1]`` and local index ``0 <-> +1``,
  ``1 <-> -1`` (the ``StaticRange(1, -2, 2)`` ordering of
  netket/hilbert/spin.py:165-178 and netket/utils/static_range.py:148-192);
* samples are stored as ``int8`` (the sampler dtype the reference derives from
  the local states, netket/sampler/base.py:122-130);
* ``total_sz`` adds the sum constraint ``sum(sigma) == 2*total_sz``
  (netket/hilbert/spin.py:187, netket/hilbert/constraint/sum.py).
"""

from __future__ import annotations

import numpy as np


class Spin:
    """Spin-s Hilbert space restricted to s = 1/2 (the only case the path uses)."""

    def __init__(self, s: float = 0.5, N: int = 1, total_sz: float | None = None):
        if float(s) != 0.5:
            raise NotImplementedError("only spin-1/2 is supported on the MI355X path")
        N = int(N)
        if N < 1:
            raise ValueError("N must be a positive integer")
        if total_sz is not None:
            two_sz = 2 * float(total_sz)
            if not two_sz.is_integer():
                raise ValueError("total_sz must be a multiple of 1/2")
            n_down2 = N - int(two_sz)
            if n_down2 % 2 != 0 or not (0 <= n_down2 // 2 <= N):
                raise ValueError(
                    f"Cannot fix the total magnetization: 2|M| must be <= N and "
                    f"have the parity of N (N={N}, total_sz={total_sz})"
                )
        self._s = 0.5
        self._N = N
        self._total_sz = None if total_sz is None else float(total_sz)

    # -- reference-compatible properties ---------------------------------
    @property
    def size(self) -> int:
        return self._N

    @property
    def local_size(self) -> int:
        return 2

    @property
    def local_states(self) -> np.ndarray:
        return np.array([1.0, -1.0])

    @property
    def shape(self) -> tuple:
        return (2,) * self._N

    @property
    def total_sz(self):
        return self._total_sz

    @property
    def constrained(self) -> bool:
        return self._total_sz is not None

    @property
    def n_down(self) -> int | None:
        """Number of sites with spin -1 (local index 1) under the Sz constraint."""
        if self._total_sz is None:
            return None
        return (self._N - int(round(2 * self._total_sz))) // 2

    @property
    def n_states(self) -> int:
        if self._total_sz is None:
            return 2**self._N
        from math import comb

        return comb(self._N, self.n_down)

    def size_at_index(self, i: int) -> int:
        return 2

    def states_to_local_indices(self, x):
        """+1 -> 0 and -1 -> 1, as ``StaticRange(1, -2, 2)``."""
        return ((1 - x) // 2) if not hasattr(x, "dtype") or x.dtype.kind in "iu" else (1 - x) / 2

    def local_indices_to_states(self, idx, dtype=np.int8):
        return (1 - 2 * np.asarray(idx)).astype(dtype)

    def all_states(self) -> np.ndarray:
        """All basis states (int8), constrained ones filtered; small N only."""
        if self._N > 24:
            raise RuntimeError("all_states is only available for N <= 24")
        n = np.arange(2**self._N, dtype=np.int64)
        bits = (n[:, None] >> np.arange(self._N - 1, -1, -1)[None, :]) & 1
        states = (1 - 2 * bits).astype(np.int8)
        if self._total_sz is not None:
            states = states[states.sum(axis=1) == int(round(2 * self._total_sz))]
        return states

    def __eq__(self, other):
        return (
            isinstance(other, Spin)
            and other._N == self._N
            and other._total_sz == self._total_sz
        )

    def __hash__(self):
        return hash(("Spin", self._N, self._total_sz))

    def __repr__(self):
        c = "" if self._total_sz is None else f", total_sz={self._total_sz}"
        return f"Spin(s=1/2{c}, N={self._N})"
