"""Metropolis samplers of the VMC path.

Mirrors ``MetropolisSampler`` (netket/sampler/metropolis.py:206-505) with the
``LocalRule`` (netket/sampler/rules/local.py:40-49) and ``ExchangeRule``
(netket/sampler/rules/exchange.py:60-218) for spin-1/2 Hilbert spaces, and the
``MetropolisLocal`` / ``MetropolisExchange`` factories
(netket/sampler/metropolis.py:532-683).

The chains live on the GPU; one call of the fused kernel performs the reset
(log_prob recomputed), the discarded sweeps and the kept sweeps that
``MCState.sample`` issues as three separate jitted calls
(netket/vqs/mc/mc_state/state.py:521-576).  Random numbers come from a
counter-based Philox4x32-10 stream keyed by the sampler seed, so chains are
reproducible from (seed, number of proposals made), not from a JAX key.
"""

from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as nat
from . import _ops
from .hilbert import Spin


class MetropolisRule:
    kind: int = -1


class LocalRule(MetropolisRule):
    """Flip one uniformly chosen site (netket/sampler/rules/local.py:40-49)."""

    kind = nat.RULE_LOCAL

    def __repr__(self):
        return "LocalRule()"


class ExchangeRule(MetropolisRule):
    """Swap the two sites of a cluster with different spins.

    Clusters come from ``clusters`` or, given a ``graph``, from all pairs at
    graph distance ``<= d_max`` in row-major order with ``i < j``
    (``compute_clusters``, netket/sampler/rules/exchange.py:190-205).

    ``probabilities`` (netket/sampler/rules/exchange.py:86-90, 116-141) gives
    each cluster (or, with ``graph``, each graph distance 1..d_max) a relative
    weight; a differing cluster is proposed with probability proportional to
    it and the move carries the log-correction log W(x) - log W(x'), W the
    total weight of the differing clusters.  The GPU sampler draws with
    integer weights, ``round(2^20 p / max p)`` (at least 1): the choice and the
    totals are then exact integer arithmetic, and the chain's stationary
    distribution stays |psi|^2 exactly, since the same weights enter the
    correction; only the proposal probabilities are rounded (relative 2^-20).
    """

    kind = nat.RULE_EXCHANGE
    WEIGHT_BITS = 20

    def __init__(self, *, clusters=None, graph=None, d_max: int = 1, probabilities=None):
        if probabilities is not None:
            probabilities = np.atleast_1d(np.asarray(probabilities, dtype=np.float64))
            if not np.all(probabilities > 0):
                raise ValueError("Probabilities must be positive")
        if clusters is None and graph is not None:
            clusters, probabilities = compute_clusters(graph, d_max, probabilities)
        elif not (clusters is not None and graph is None):
            raise ValueError("You must either provide the list of exchange-clusters or a netket graph")
        self.clusters = np.asarray(clusters, dtype=np.int32).reshape(-1, 2)
        if probabilities is not None and len(probabilities) != len(self.clusters):
            raise TypeError("Number of clusters and probabilities don't match: "
                            f"{len(self.clusters)} != {len(probabilities)}")
        self.probabilities = probabilities
        self.weights = None if probabilities is None else integer_weights(probabilities, self.WEIGHT_BITS)
        self._dev = {}

    def device_clusters(self):
        return self._device("clusters", self.clusters)

    def device_weights(self):
        """The integer cluster weights on the current device (None: uniform)."""
        return None if self.weights is None else self._device("weights", self.weights.view(np.int32))

    def _device(self, name, arr):
        dev = _ops.device()
        t = self._dev.get((name, dev))
        if t is None:
            t = torch.as_tensor(arr, device=dev).contiguous()
            self._dev[(name, dev)] = t
        return t

    def __repr__(self):
        return f"ExchangeRule(# of clusters: {len(self.clusters)})"


def integer_weights(p, bits: int = 20) -> np.ndarray:
    """round(2^bits p / max p), at least 1, as uint32."""
    p = np.asarray(p, dtype=np.float64)
    w = np.rint(p / p.max() * float(1 << bits))
    return np.maximum(w, 1.0).astype(np.uint32)


def compute_clusters(graph, d_max: int, prob=None):
    """(clusters, per-cluster probabilities or None); ``prob`` has one entry per
    graph distance 1..d_max (netket/sampler/rules/exchange.py:190-205)."""
    distances = np.asarray(graph.distances())
    clusters = np.argwhere(distances <= d_max)
    clusters = clusters[clusters[:, 0] < clusters[:, 1]]
    if prob is not None:
        prob = np.asarray(prob, dtype=np.float64)
        if prob.shape != (d_max,):
            raise ValueError(f"Expected d_max = {d_max} probabilities, got {prob}")
        prob = prob[distances[clusters[:, 0], clusters[:, 1]].astype(np.int64) - 1]
    return clusters.astype(np.int32), prob


@dataclass
class MetropolisSamplerState:
    """Chain state on the device (netket/sampler/metropolis.py:42-137)."""

    sigma: torch.Tensor                  # (n_chains, N) int8
    log_prob: torch.Tensor | None        # (n_chains,) model dtype
    seed: int
    counter: int = 0                     # proposals drawn so far per chain
    n_accepted_proc: torch.Tensor | None = None   # (n_chains,) int64
    n_steps_proc: int = 0
    counter_dev: torch.Tensor | None = None  # the counter on the device (source of truth)
    extra: dict = field(default_factory=dict)

    @property
    def acceptance(self) -> float | None:
        if self.n_steps_proc == 0 or self.n_accepted_proc is None:
            return None
        return float(self.n_accepted_proc.sum().item()) / self.n_steps_proc

    @property
    def n_accepted(self) -> int:
        return int(self.n_accepted_proc.sum().item())

    @property
    def n_steps(self) -> int:
        return self.n_steps_proc


class MetropolisSampler:
    is_exact = False

    def __init__(self, hilbert: Spin, rule: MetropolisRule, *, sweep_size: int | None = None,
                 reset_chains: bool = False, n_chains: int | None = None,
                 n_chains_per_rank: int | None = None, chunk_size: int | None = None,
                 machine_pow: int = 2, dtype=np.int8):
        if not isinstance(rule, MetropolisRule):
            raise TypeError("The second positional argument, rule, must be a MetropolisRule")
        if not isinstance(reset_chains, bool):
            raise TypeError("reset_chains must be a boolean.")
        if n_chains is not None and n_chains_per_rank is not None:
            raise ValueError("Cannot specify both `n_chains` and `n_chains_per_rank`")
        if n_chains is None:
            n_chains = 16 if n_chains_per_rank is None else n_chains_per_rank
        if sweep_size is None:
            sweep_size = hilbert.size
        if np.dtype(dtype) != np.int8:
            warnings.warn("samples are produced as int8 spin values on the GPU path", stacklevel=2)
        if isinstance(rule, ExchangeRule) and rule.clusters.size and rule.clusters.max() >= hilbert.size:
            raise ValueError("exchange cluster refers to a site outside the Hilbert space")
        self.hilbert = hilbert
        self.rule = rule
        self.sweep_size = int(sweep_size)
        self.reset_chains = reset_chains
        self.n_chains = int(n_chains)
        self.chunk_size = chunk_size
        self.machine_pow = machine_pow
        self.dtype = np.int8

    @property
    def n_batches(self) -> int:
        return self.n_chains

    @property
    def n_chains_per_rank(self) -> int:
        return self.n_chains

    def _random_sigma(self, seed: int, counter: int) -> torch.Tensor:
        N = self.hilbert.size
        sigma = torch.empty((self.n_chains, N), dtype=torch.int8, device=_ops.device())
        n_down = self.hilbert.n_down if self.hilbert.constrained else -1
        nat.call("nka_random_states", _ops.ptr(sigma), self.n_chains, N, n_down, seed & (2**64 - 1),
                 counter, _ops.stream())
        return sigma

    def init_state(self, machine, parameters, seed: int | None = None) -> MetropolisSamplerState:
        """Random initial chains (netket/sampler/metropolis.py:353-380)."""
        if seed is None:
            seed = int(np.random.SeedSequence().generate_state(1, np.uint64)[0])
        seed = int(seed)
        sigma = self._random_sigma(seed, 0)
        nacc = torch.zeros(self.n_chains, dtype=torch.int64, device=sigma.device)
        counter = 1 << 40
        cdev = torch.tensor([counter], dtype=torch.int64, device=sigma.device)
        return MetropolisSamplerState(sigma=sigma, log_prob=None, seed=seed, counter=counter,
                                      n_accepted_proc=nacc, counter_dev=cdev)

    def reset(self, machine, parameters, state: MetropolisSamplerState) -> MetropolisSamplerState:
        """Zero the acceptance counters; redraw the chains when reset_chains is set."""
        if self.reset_chains:
            state.sigma = self._random_sigma(state.seed, state.counter)
            state.counter += 1 << 20
        state.n_accepted_proc.zero_()
        state.n_steps_proc = 0
        return state

    def _check_machine(self, machine):
        from .models import FFNN, RBM

        if not isinstance(machine, (RBM, FFNN)):
            raise NotImplementedError(
                f"the GPU Metropolis kernels support the RBM and FFNN ansatze, not {type(machine).__name__}")

    def sample_fused(self, machine, parameters: torch.Tensor, state: MetropolisSamplerState,
                     n_discard: int, chain_length: int, return_log_probabilities: bool = False):
        """Reset log_prob, run ``n_discard`` discarded and ``chain_length`` kept sweeps.

        Returns samples of shape (n_chains, chain_length, N) (and their
        log-probabilities when requested) and the updated state.
        """
        self._check_machine(machine)
        N = self.hilbert.size
        dev = state.sigma.device
        mt = machine.torch_dtype()
        if state.log_prob is None or state.log_prob.dtype != mt:
            state.log_prob = torch.empty(self.n_chains, dtype=mt, device=dev)
        samples = torch.empty((self.n_chains, chain_length, N), dtype=torch.int8, device=dev)
        lps = (torch.empty((self.n_chains, chain_length), dtype=mt, device=dev)
               if return_log_probabilities else None)
        m = machine.native(parameters, N)
        w_ptr = C.c_void_p(0)
        if isinstance(self.rule, ExchangeRule):
            cl = self.rule.device_clusters()
            cl_ptr, n_cl = _ops.ptr(cl), int(cl.shape[0])
            wt = self.rule.device_weights()
            if wt is not None:
                w_ptr = _ops.ptr(wt)
        else:
            cl_ptr, n_cl = C.c_void_p(0), 0
        with _ops.TIMER.region("metropolis"):
            nat.call(machine.metropolis_symbol, C.byref(m), self.rule.kind, cl_ptr, n_cl, w_ptr,
                     _ops.ptr(state.sigma), _ops.ptr(state.log_prob), self.n_chains,
                     self.sweep_size, int(n_discard), int(chain_length), float(self.machine_pow),
                     state.seed & (2**64 - 1), state.counter, _ops.ptr(state.counter_dev),
                     _ops.ptr(samples), _ops.ptr(lps), _ops.ptr(state.n_accepted_proc),
                     _ops.stream())
        n_prop = (int(n_discard) + int(chain_length)) * self.sweep_size
        state.counter += n_prop
        state.n_steps_proc += n_prop * self.n_chains
        if return_log_probabilities:
            return (samples, lps), state
        return samples, state

    def sample(self, machine, parameters, *, state=None, chain_length: int = 1,
               return_log_probabilities: bool = False):
        """``Sampler.sample`` (netket/sampler/base.py:344-383)."""
        if state is None:
            state = self.init_state(machine, parameters)
            state = self.reset(machine, parameters, state)
        return self.sample_fused(machine, parameters, state, 0, chain_length,
                                 return_log_probabilities)

    def __repr__(self):
        return (f"MetropolisSampler(rule = {self.rule}, n_chains = {self.n_chains}, "
                f"sweep_size = {self.sweep_size}, reset_chains = {self.reset_chains}, "
                f"machine_power = {self.machine_pow}, dtype = int8)")


def MetropolisLocal(hilbert, **kwargs) -> MetropolisSampler:
    """Single-spin-flip Metropolis sampler (netket/sampler/metropolis.py:532-572)."""
    return MetropolisSampler(hilbert, LocalRule(), **kwargs)


def MetropolisExchange(hilbert, *, clusters=None, graph=None, d_max: int = 1, probabilities=None,
                       **kwargs) -> MetropolisSampler:
    """Exchange Metropolis sampler (netket/sampler/metropolis.py:575-683)."""
    rule = ExchangeRule(clusters=clusters, graph=graph, d_max=d_max, probabilities=probabilities)
    return MetropolisSampler(hilbert, rule, **kwargs)
