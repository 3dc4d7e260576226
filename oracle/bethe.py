"""Ground-state energy of the periodic spin-1/2 Heisenberg chain from the
Bethe equations — oracle (test infrastructure only).

The reference states converged energies against exact diagonalisation
(netket/exact.py:65-97, lanczos_ed).  Beyond the sizes ED reaches, the chain's
ground state is fixed by the Bethe ansatz (H. Bethe, Z. Phys. 71, 205 (1931);
the logarithmic form as in M. Karbach and G. Mueller, Computers in Physics 11,
36 (1997)).  With M = N/2 rapidities lambda_j and the ground-state quantum
numbers I_j = -(M-1)/2, ..., (M-1)/2:

    N * 2 atan(2 lambda_j) = 2 pi I_j + sum_{k != j} 2 atan(lambda_j - lambda_k)

and, for H = J sum_i S_i . S_{i+1} (spin operators, J > 0),

    E = J N / 4 - J sum_j 2 / (4 lambda_j^2 + 1).

The reference's Heisenberg operator uses Pauli matrices
(netket/operator/_heisenberg.py: sigma^z sigma^z + 2 (sigma^+ sigma^- + h.c.)),
four times the spin operators, so its energy is 4 E.  Pinned in the tests by
exact diagonalisation of small chains and the L = 22 value of
docs/tutorials/gs-heisenberg.ipynb:163.
"""

from __future__ import annotations

import numpy as np


def rapidities(N: int, tol: float = 1e-15, max_iter: int = 100000) -> np.ndarray:
    """Ground-state rapidities of the N-site periodic chain (N even, M = N/2)."""
    if N % 2:
        raise ValueError("the Sz = 0 ground state needs an even number of sites")
    M = N // 2
    I = np.arange(M) - (M - 1) / 2.0
    lam = 0.5 * np.tan(np.pi * I / N)
    for _ in range(max_iter):
        d = lam[:, None] - lam[None, :]
        phase = (np.pi * I + np.sum(np.arctan(d), axis=1)) / N
        new = 0.5 * np.tan(phase)
        if np.max(np.abs(new - lam)) < tol:
            return new
        # damped update: the plain fixed point converges, damping makes it robust
        lam = 0.5 * (lam + new)
    raise RuntimeError("Bethe equations did not converge")


def heisenberg_e0(N: int, J: float = 1.0) -> float:
    """E0 of the reference's Heisenberg chain (Pauli matrices), periodic, N sites."""
    lam = rapidities(N)
    e_spin = J * N / 4.0 - J * np.sum(2.0 / (4.0 * lam**2 + 1.0))
    return float(4.0 * e_spin)
