"""Monte Carlo variational state (netket/vqs/mc/mc_state/state.py:180-1016).

``MCState`` owns the parameters (one flat device buffer), the sampler state and
the cached samples, and computes local estimators, expectation values and
the energy gradient on the GPU:

* ``sample``            netket/vqs/mc/mc_state/state.py:521-576
* ``local_estimators``  netket/vqs/mc/mc_state/local_estimators.py:34-68
* ``expect``            netket/vqs/mc/mc_state/expect.py
* ``expect_and_grad``   expect_and_forces + force_to_grad
                        (netket/vqs/mc/mc_state/expect_forces.py:39-112,
                         netket/vqs/mc/common.py:102-118)
"""

from __future__ import annotations

import math
import warnings

import numpy as np
import torch

from . import _engine as eng
from . import _ops
from .stats import Stats, statistics


def compute_chain_length(n_chains: int, n_samples: int) -> int:
    """netket/vqs/mc/mc_state/state.py:60-79."""
    if n_samples <= 0:
        raise ValueError(f"Invalid number of samples: n_samples={n_samples}")
    chain_length = int(np.ceil(n_samples / n_chains))
    if chain_length * n_chains != n_samples:
        warnings.warn(
            f"n_samples={n_samples} does not divide n_chains={n_chains}, increased to "
            f"{chain_length * n_chains}", UserWarning, stacklevel=3)
    return chain_length


class MCState:
    def __init__(self, sampler, model=None, *, n_samples: int | None = None,
                 n_samples_per_rank: int | None = None, n_discard_per_chain: int | None = None,
                 chain_length: int | None = None, seed: int | None = None,
                 sampler_seed: int | None = None, variables=None, chunk_size: int | None = None):
        self.sampler = sampler
        self.model = model
        self.hilbert = sampler.hilbert
        N = self.hilbert.size
        if seed is None:
            seed = int(np.random.SeedSequence().generate_state(1, np.uint64)[0] >> 1)
        if sampler_seed is None:
            sampler_seed = (int(seed) * 0x9E3779B97F4A7C15 + 1) & (2**63 - 1)
        self._seed = int(seed)
        if variables is not None:
            params = variables["params"] if "params" in variables else variables
            self._flat = model.flatten(params, N)
        else:
            self._flat = model.init(self._seed, N)
        self.sampler_state = sampler.init_state(model, self._flat, sampler_seed)
        self.sampler_state = sampler.reset(model, self._flat, self.sampler_state)
        if n_samples is not None and n_samples_per_rank is not None:
            raise ValueError("Only one argument between `n_samples` and `n_samples_per_rank`")
        if n_samples_per_rank is not None:
            n_samples = n_samples_per_rank
        if chain_length is not None and n_samples is not None:
            raise ValueError("Cannot specify both `chain_length` and `n_samples`.")
        if chain_length is None:
            chain_length = compute_chain_length(sampler.n_chains, 1000 if n_samples is None else n_samples)
        self._chain_length = int(chain_length)
        self.n_discard_per_chain = n_discard_per_chain
        self._chunk_size = None
        self.chunk_size = chunk_size
        self._samples = None
        self._jac_cache = None

    # -- sizes ---------------------------------------------------------------
    @property
    def chain_length(self) -> int:
        return self._chain_length

    @chain_length.setter
    def chain_length(self, v: int):
        self._chain_length = int(v)
        self.reset()

    @property
    def n_samples(self) -> int:
        return self._chain_length * self.sampler.n_chains

    @n_samples.setter
    def n_samples(self, n: int):
        self._chain_length = compute_chain_length(self.sampler.n_chains, n)
        self.reset()

    @property
    def n_discard_per_chain(self) -> int:
        return self._n_discard

    @n_discard_per_chain.setter
    def n_discard_per_chain(self, v):
        if v is not None and v < 0:
            raise ValueError(f"Invalid number of discarded samples: n_discard_per_chain={v}")
        self._n_discard = 5 if v is None else int(v)

    @property
    def chunk_size(self) -> int | None:
        """Largest number of samples whose local energies are evaluated in one
        launch (netket/vqs/mc/mc_state/state.py:471-512); None evaluates all at
        once.  The values do not depend on it."""
        return self._chunk_size

    @chunk_size.setter
    def chunk_size(self, chunk_size):
        if chunk_size is None:
            self._chunk_size = None
            return
        if not isinstance(chunk_size, (int, np.integer)) or isinstance(chunk_size, bool) or chunk_size <= 0:
            raise ValueError(f"Chunk size must be a positive INTEGER (got {chunk_size} instead).")
        if chunk_size & (chunk_size - 1):
            warnings.warn("For performance reasons, we suggest to use a power-of-two chunk size.",
                          stacklevel=2)
        self._chunk_size = int(chunk_size)

    @property
    def n_parameters(self) -> int:
        return self._flat.numel()

    # -- parameters ------------------------------------------------------------
    @property
    def parameters_flat(self) -> torch.Tensor:
        return self._flat

    @property
    def parameters(self) -> dict:
        return self.model.views(self._flat, self.hilbert.size)

    @parameters.setter
    def parameters(self, params):
        if isinstance(params, torch.Tensor) and params.ndim == 1:
            self._flat.copy_(params.to(self._flat.dtype))
        else:
            self._flat.copy_(self.model.flatten(params, self.hilbert.size))
        self.reset()

    @property
    def variables(self) -> dict:
        return {"params": self.parameters}

    def flatten_tree(self, tree) -> torch.Tensor:
        """A parameter-shaped dict (or flat tensor) as a flat FP64 device vector."""
        if isinstance(tree, torch.Tensor):
            return tree.to(device=self._flat.device, dtype=torch.float64).reshape(-1)
        return self.model.flatten(tree, self.hilbert.size).to(torch.float64)

    def unflatten(self, flat: torch.Tensor) -> dict:
        return self.model.views(flat.to(self._flat.dtype).contiguous(), self.hilbert.size)

    # -- sampling ----------------------------------------------------------------
    def reset(self):
        self._samples = None
        self._jac_cache = None

    def sample(self, *, chain_length: int | None = None, n_samples: int | None = None,
               n_discard_per_chain: int | None = None) -> torch.Tensor:
        if n_samples is None and chain_length is None:
            chain_length = self._chain_length
        elif chain_length is None:
            chain_length = compute_chain_length(self.sampler.n_chains, n_samples)
        if n_discard_per_chain is None:
            n_discard_per_chain = self._n_discard
        self.sampler_state = self.sampler.reset(self.model, self._flat, self.sampler_state)
        self._samples, self.sampler_state = self.sampler.sample_fused(
            self.model, self._flat, self.sampler_state, n_discard_per_chain, chain_length)
        self._jac_cache = None
        return self._samples

    @property
    def samples(self) -> torch.Tensor:
        """Cached samples (n_chains, chain_length, N), sampled on first access."""
        if self._samples is None:
            self.sample()
        return self._samples

    def log_value(self, sigma) -> torch.Tensor:
        return self.model.apply(self._flat, sigma)

    # -- estimators ----------------------------------------------------------------
    def local_estimators(self, op, *, chunk_size: int | None = None) -> torch.Tensor:
        """E_loc of the cached samples, shape (n_chains, chain_length), FP64, in
        launches of at most ``chunk_size`` (default: the state's) samples."""
        s = self.samples
        C, L, N = s.shape
        flat_s = s.reshape(C * L, N)
        cs = self._chunk_size if chunk_size is None else int(chunk_size)
        if cs is None or cs >= C * L:
            e = eng.local_energies(self.model, self._flat, op, flat_s)
        else:
            e = torch.empty(C * L, dtype=torch.float64, device=s.device)
            for i in range(0, C * L, cs):
                e[i:i + cs] = eng.local_energies(self.model, self._flat, op, flat_s[i:i + cs])
        return e.reshape(C, L)

    def expect(self, op) -> Stats:
        return statistics(self.local_estimators(op))

    def _prepared(self):
        """tanh(theta) of the cached samples, kept until the samples change."""
        if self._jac_cache is None or self._jac_cache.get("samples") is not self._samples:
            s = self.samples
            C, L, N = s.shape
            sig = s.reshape(C * L, N)
            t, _ = eng.prepare(self.model, self._flat, sig, row_major=True)
            self._jac_cache = {"samples": self._samples, "sigma": sig, "t": t}
        return self._jac_cache

    def expect_and_forces_flat(self, op):
        """(Stats, 2 Re F as a flat FP64 device vector, E_loc (Ns,))."""
        eloc = self.local_estimators(op)
        st = statistics(eloc)
        pc = self._prepared()
        ns = eloc.numel()
        w = eng.centre_scale(eloc.reshape(-1), st.mean_device, 2.0 / ns)
        g = eng.contract(self.model, self._flat, pc["sigma"], pc["t"], w)[0]
        return st, g, eloc

    def expect_and_grad(self, op, *, mutable=None, use_covariance=None):
        """(Stats, gradient dict in the parameter dtype): 2 Re Cov(O, E_loc)."""
        st, g, _ = self.expect_and_forces_flat(op)
        grad_flat = g.to(self._flat.dtype)
        self._last_grad64 = g
        return st, self.model.views(grad_flat, self.hilbert.size)

    def expect_and_forces(self, op):
        st, g, _ = self.expect_and_forces_flat(op)
        return st, self.model.views((0.5 * g).to(self._flat.dtype), self.hilbert.size)

    def quantum_geometric_tensor(self, qgt_T=None, **kwargs):
        from .optimizer.qgt import QGTJacobianDense

        qgt_T = qgt_T or QGTJacobianDense
        return qgt_T(self, **kwargs)

    def __repr__(self):
        return (f"MCState(hilbert = {self.hilbert}, sampler = {self.sampler}, "
                f"n_samples = {self.n_samples}, n_discard_per_chain = {self._n_discard}, "
                f"n_parameters = {self.n_parameters})")


def _sqrt(n):
    return math.sqrt(n)
