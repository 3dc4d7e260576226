"""CPU FP64 oracle for the VMC step — TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the reference algorithm of NetKet's VMC
step (the functions listed in each module's docstring, with file:line of the
reference).  It is the checker of the GPU path and nothing else: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and the product (``netket_amd``) never does.

How each part is pinned (see ``tests/test_oracle_*.py``):

* philox       : Random123 known-answer vectors of Philox4x32-10.
* operators    : Ising conn vs a Kronecker-product Hamiltonian; exact ground
                 energies (TFIM closed form for L=8/16/20, netket/exact.py:61-63,
                 docs/tutorials/gs-ising.ipynb:282; Heisenberg L=22 Sz=0
                 -39.147522607062776, docs/tutorials/gs-heisenberg.ipynb:163)
                 through Lanczos on the oracle's sparse H; LocalOperator conn
                 vs the Kronecker Hamiltonian elementwise.
* rbm          : log psi vs an independent dense evaluation; E_loc vs
                 (H psi)(s)/psi(s) from the dense H; derivatives vs central
                 finite differences; the energy gradient (forces_to_grad)
                 and the SR right-hand side dv vs central differences of the
                 exact energy <psi|H|psi>/<psi|psi> (test_oracle_gradient).
* stats        : numpy identities and hand-computed branch cases.
* ngd          : SR == minSR (Woodbury identity, test/driver/test_srt.py:98-211),
                 scipy cho_solve, CG residual bound, pinv_smooth vs lstsq.
* sampler      : exact |psi|^2 of small systems; detailed balance of the
                 exchange correction.
"""
