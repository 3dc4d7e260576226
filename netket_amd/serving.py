"""Continuous batching of independent ground-state VMC jobs on one GPU.

NetKet runs one driver loop per process (netket/_src/driver/abstract_variational_driver.py:349-499);
this module runs many such loops on one device.  A request stream (Poisson
arrivals of transverse-field Ising jobs of uneven size) is admitted by HBM
footprint; every admitted job keeps its parameters, chains and solver
buffers resident on the GPU.

Jobs admitted together that have the same shape (lattice, hidden units,
samples) form a group whose Metropolis sweeps run as one shared launch
(``nka_rbm_metropolis_batched``: the job index in the grid) and whose local
energies run as one shared launch (``nka_rbm_local_energy_batched``, when the
model's W fits the LDS; otherwise one launch per job); the rest of each job's
step (statistics, Jacobian, S or T, Cholesky, update) follows on the same
stream.  A group's step is captured once into a HIP graph.  Groups progress
independently: the host polls each group's completion event and enqueues its
next step as soon as its previous one has finished, so a large job never
gates a small one; step latencies are the GPU times between events recorded
around each step.  A job's result depends only on its own seeds and shape:
its Philox counter lives in its own device state and the shared launches do
per job exactly what a group of one does, so it is the same alone and in
the mix.
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import graph as nkg
from . import hilbert as nkh
from . import models as nkm
from . import operator as nko
from . import sampler as nks
from . import vqs as nkv
from . import _native as nat
from . import _ops
from .driver import VMC_SR, statistics_from_moments
from .optimizer import Sgd
from .optimizer import solver as nksolver


@dataclass
class JobSpec:
    job_id: int
    arrival: float          # seconds after the start of the stream
    dim: int                # 1: chain of L sites, 2: L x L square lattice
    L: int
    h: float
    alpha: int
    n_samples: int
    n_steps: int = 200
    J: float = 1.0
    seed: int = 0
    lr: float = 0.02
    diag_shift: float = 0.01

    @property
    def n_sites(self) -> int:
        return self.L if self.dim == 1 else self.L * self.L


def make_stream(n_jobs: int = 256, rate: float = 4.0, seed: int = 0, n_steps: int = 200):
    """The request stream of the serving workload: Poisson arrivals at ``rate``
    jobs/s; each job a TFIM, 1-D with L uniform in {16..128} or 2-D with side
    uniform in {4..12} (equal odds), h uniform in [0.5, 2], J = 1, FP32 RBM
    with alpha in {1, 2, 4}, samples per step in {1024, 4096, 8192}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for j in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        dim = 1 if rng.random() < 0.5 else 2
        L = int(rng.integers(16, 129)) if dim == 1 else int(rng.integers(4, 13))
        out.append(JobSpec(job_id=j, arrival=t, dim=dim, L=L, h=float(rng.uniform(0.5, 2.0)),
                           alpha=int(rng.choice([1, 2, 4])),
                           n_samples=int(rng.choice([1024, 4096, 8192])), n_steps=n_steps,
                           seed=int(rng.integers(0, 2**31 - 1))))
    return out


def footprint_bytes(spec: JobSpec) -> int:
    """Device memory of one job: parameters, chains, samples, tanh tables,
    the operands of S (spin pairs and activation pairs hi/lo with their
    product) or of T (centred Jacobian bf16 hi/lo), the FP64 S or T matrix and
    workspaces, doubled for the graph pool kept next to the eager warm-up
    buffers."""
    N = spec.n_sites
    M = spec.alpha * N
    Np = M + N * M + N
    Ns = spec.n_samples
    ntk = Np > Ns
    n = Ns if ntk else Np
    ld = ((Ns if not ntk else Np) + 63) // 64 * 64
    rows = Np if not ntk else Ns
    b = 4 * Np * 4                        # parameters, gradient, update, means (f32/f64)
    b += Ns * N * 2                       # samples + chain state
    b += Ns * M * 4 * 2                   # t and its transpose
    if ntk:
        b += rows * ld * 4                # X hi + lo
    else:
        na, nb = 1 + N + N * (N - 1) // 2, 1 + M + M * (M + 1) // 2
        b += N * ld + na * ld * 2 + nb * ld * 4 + na * ((nb + 63) // 64 * 64) * 4  # spins^T, P, Q, C
    b += n * n * 8                        # S or T
    b += (Ns // 256 + 1) * (Np + 1) * 8   # contraction partials
    b += n * 16 + (n // 64 + 1) * 64 * 64 * 8
    return int(2 * b) + (32 << 20)


def tfim_chain_exact_energy(L: int, h: float, J: float = 1.0) -> float:
    """Ground-state energy of the periodic transverse-field Ising chain
    H = J sum z_i z_{i+1} - h sum x_i (netket/operator/_ising/jax.py:125-165) from
    its free-fermion solution: for even L (or J < 0) the antiperiodic modes
    k = pi (2n+1) / L; for odd L with J > 0 (a frustrated ring) the periodic
    modes k = 2 pi n / L with the k = 0 mode entering as -(h - |J|)."""
    Ja = abs(J)
    eps = lambda k: np.sqrt(Ja * Ja + h * h - 2.0 * Ja * h * np.cos(k))
    n = np.arange(L)
    if L % 2 == 0 or J < 0:
        return float(-np.sum(eps(np.pi * (2 * n + 1) / L)))
    return float(-np.sum(eps(2.0 * np.pi * n[1:] / L)) - (h - Ja))


class Job:
    """One job's state: Hamiltonian, RBM state and VMC_SR driver, plus the
    persistent buffers the group's shared launches write into."""

    def __init__(self, spec: JobSpec, n_chains: int | None = None):
        self.spec = spec
        N = spec.n_sites
        g = nkg.Chain(spec.L) if spec.dim == 1 else nkg.Square(spec.L)
        hi = nkh.Spin(0.5, N=N)
        self.H = nko.Ising(hi, g, h=spec.h, J=spec.J)
        if n_chains is None:
            n_chains = min(spec.n_samples, 1024)
        sa = nks.MetropolisLocal(hi, n_chains=n_chains)
        model = nkm.RBM(alpha=spec.alpha, param_dtype=np.float32)
        self.vs = nkv.MCState(sa, model, n_samples=spec.n_samples, n_discard_per_chain=5,
                              seed=spec.seed)
        self.drv = VMC_SR(self.H, Sgd(spec.lr), diag_shift=spec.diag_shift,
                          variational_state=self.vs, linear_solver=nksolver.cholesky)
        vs = self.vs
        st = vs.sampler_state
        C, Lc = sa.n_chains, vs.chain_length
        dev = st.sigma.device
        st.log_prob = torch.empty(C, dtype=model.torch_dtype(), device=dev)
        self.samples = torch.empty((C, Lc, N), dtype=torch.int8, device=dev)
        self.eloc = torch.empty((C, Lc), dtype=torch.float64, device=dev)
        self.steps_done = 0
        self.stats = []

    @property
    def shape_key(self):
        s = self.spec
        return (s.dim, s.L, s.alpha, s.n_samples, self.vs.sampler.n_chains)

    @property
    def done(self) -> bool:
        return self.steps_done >= self.spec.n_steps

    def release(self):
        self.drv = None
        self.vs = None
        self.samples = None
        self.eloc = None


class JobGroup:
    """Jobs of one shape stepped together (see the module docstring)."""

    def __init__(self, jobs, stream):
        self.jobs = list(jobs)
        self.stream = stream
        j0 = self.jobs[0]
        self.N = j0.spec.n_sites
        vs0 = j0.vs
        self.model = vs0.model
        self.shape = self.model.native(vs0.parameters_flat, self.N)
        sa = vs0.sampler
        self.C, self.Lc, self.D = sa.n_chains, vs0.chain_length, vs0._n_discard
        self.sweep_size = sa.sweep_size
        self.machine_pow = float(sa.machine_pow)
        dev = vs0.parameters_flat.device
        n = len(self.jobs)
        sweep = (nat.NkaSweepJob * n)()
        for k, j in enumerate(self.jobs):
            st = j.vs.sampler_state
            sweep[k].model = j.vs.model.native(j.vs.parameters_flat, self.N)
            sweep[k].sigma = st.sigma.data_ptr()
            sweep[k].log_prob = st.log_prob.data_ptr()
            sweep[k].samples = j.samples.data_ptr()
            sweep[k].log_prob_samples = None
            sweep[k].n_accepted = st.n_accepted_proc.data_ptr()
            sweep[k].seed = st.seed & (2**64 - 1)
            sweep[k].counter = st.counter
            sweep[k].counter_dev = st.counter_dev.data_ptr()
        self._sweep_dev = torch.frombuffer(bytearray(bytes(sweep)), dtype=torch.uint8).to(dev)
        kind, op_ptr, self._op_keep = _ops.op_struct(j0.H)
        self.batched_eloc = False
        if kind == nat.OP_ISING:
            ising0 = C.cast(op_ptr, C.POINTER(nat.NkaIsing)).contents
            self._ising_shape = nat.NkaIsing.from_buffer_copy(ising0)
            if nat.lib().nka_rbm_local_energy_batchable(C.byref(self.shape), C.byref(self._ising_shape)):
                el = (nat.NkaElocJob * n)()
                self._op_keeps = []
                for k, j in enumerate(self.jobs):
                    kk, pp, keep = _ops.op_struct(j.H)
                    self._op_keeps.append(keep)
                    el[k].model = j.vs.model.native(j.vs.parameters_flat, self.N)
                    el[k].op = nat.NkaIsing.from_buffer_copy(C.cast(pp, C.POINTER(nat.NkaIsing)).contents)
                    el[k].sigma = j.samples.data_ptr()
                    el[k].eloc = j.eloc.data_ptr()
                    el[k].log_psi = None
                self._eloc_dev = torch.frombuffer(bytearray(bytes(el)), dtype=torch.uint8).to(dev)
                self.batched_eloc = True
        self.graph = None
        self.ws = _ops.Workspace()
        self.inflight = None
        self.latencies_ms = []

    def _step_body(self):
        """One step of every member: shared sweeps, shared (or per-job) local
        energies, then each job's solve and update."""
        n = len(self.jobs)
        for j in self.jobs:
            j.vs.sampler_state.n_accepted_proc.zero_()
        with _ops.TIMER.region("metropolis"):
            nat.call("nka_rbm_metropolis_batched", _ops.ptr(self._sweep_dev), n, C.byref(self.shape),
                     nat.RULE_LOCAL, None, 0, self.C, self.sweep_size, self.D, self.Lc, self.machine_pow,
                     _ops.stream())
        n_prop = (self.D + self.Lc) * self.sweep_size
        for j in self.jobs:
            st = j.vs.sampler_state
            st.counter += n_prop
            st.n_steps_proc = n_prop * self.C
            j.vs.reset()
            j.vs._samples = j.samples
        if self.batched_eloc:
            with _ops.TIMER.region("local_energy"):
                nat.call("nka_rbm_local_energy_batched", _ops.ptr(self._eloc_dev), n, C.byref(self.shape),
                         C.byref(self._ising_shape), self.C * self.Lc, _ops.stream())
        else:
            for j in self.jobs:
                j.eloc.copy_(j.vs.local_estimators(j.H))
        out = []
        for j in self.jobs:
            st, dp = j.drv.compute_loss_and_update(eloc=j.eloc)
            j.drv.update_parameters(dp)
            j.drv._step_count += 1
            out.append(st)
        return out

    def capture(self, warmup: int = 2):
        """``warmup`` eager steps (so every buffer exists), then one step
        captured into a HIP graph; the warm-up steps count as steps."""
        cs = torch.cuda.current_stream()
        self.stream.wait_stream(cs)
        with _ops.workspace_scope(self.ws):
            with torch.cuda.stream(self.stream):
                for _ in range(warmup):
                    sts = self._step_body()
                    for j, st in zip(self.jobs, sts):
                        j.stats.append(statistics_from_moments(st))
                        j.steps_done += 1
            self.stream.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self.stream):
                self._graph_stats = self._step_body()
        self.graph = g

    def launch(self):
        """Enqueue one captured step with events around it."""
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(self.stream):
            s.record()
            self.graph.replay()
            for j, st in zip(self.jobs, self._graph_stats):
                j.stats.append(statistics_from_moments(st))
            e.record()
        self.inflight = (s, e)

    def poll(self) -> bool:
        """True once the step in flight has finished (its latency recorded)."""
        if self.inflight is None:
            return True
        s, e = self.inflight
        if not e.query():
            return False
        self.latencies_ms.append(s.elapsed_time(e))
        for j in self.jobs:
            j.steps_done += 1
        self.inflight = None
        return True

    @property
    def done(self) -> bool:
        return all(j.done for j in self.jobs)


@dataclass
class ServingReport:
    n_jobs: int
    wall_s: float
    samples_per_s: float
    steps_per_s: float
    step_latency_ms_p50: float
    step_latency_ms_p99: float
    job_latency_s_p50: float
    job_latency_s_p99: float
    max_concurrent: int
    n_groups: int = 0
    shared_launch_jobs: int = 0
    final_energies: dict = field(default_factory=dict)
    final_params: dict = field(default_factory=dict)
    rejected: list = field(default_factory=list)
    accuracy: list = field(default_factory=list)


class Server:
    def __init__(self, hbm_budget_gb: float = 200.0, n_streams: int = 8, use_graph: bool = True,
                 max_active: int = 64, keep_params: bool = False, max_group: int = 16, batch: bool = True):
        if not use_graph:
            raise ValueError("the serving loop steps captured graphs (use_graph=True)")
        self.keep_params = keep_params
        self.budget = int(hbm_budget_gb * (1 << 30))
        self.n_streams = n_streams
        self.streams = None
        self.max_active = max_active
        self.max_group = max_group if batch else 1

    def run(self, specs, time_scale: float = 1.0, on_admit=None, on_progress=None,
            progress_every_s: float = 10.0) -> ServingReport:
        # a job larger than the whole budget can never be admitted: it is
        # reported as rejected instead of blocking the queue
        rejected = [s.job_id for s in specs if footprint_bytes(s) > self.budget]
        pending = sorted((s for s in specs if footprint_bytes(s) <= self.budget), key=lambda s: s.arrival)
        if pending and self.streams is None:
            self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        groups = []
        used = 0
        n_active = 0
        t0 = time.perf_counter()
        finished_at = {}
        step_lat = []
        total_samples = 0
        total_steps = 0
        max_conc = 0
        finals, params, accuracy = {}, {}, []
        k = 0
        n_groups = 0
        shared = 0
        last_report = time.perf_counter()
        while pending or groups:
            if on_progress is not None and time.perf_counter() - last_report >= progress_every_s:
                last_report = time.perf_counter()
                on_progress(len(finished_at), n_active, len(pending))
            now = time.perf_counter() - t0
            # admission: every arrived job that fits, grouped by shape
            admitted = []
            while (pending and pending[0].arrival * time_scale <= now and n_active + len(admitted) < self.max_active
                   and used + footprint_bytes(pending[0]) <= self.budget):
                spec = pending.pop(0)
                used += footprint_bytes(spec)
                admitted.append(spec)
            by_key = {}
            for spec in admitted:
                by_key.setdefault((spec.dim, spec.L, spec.alpha, spec.n_samples, spec.n_steps), []).append(spec)
            for key, members in by_key.items():
                for i in range(0, len(members), self.max_group):
                    chunk = members[i:i + self.max_group]
                    stream = self.streams[k % len(self.streams)]
                    k += 1
                    # the jobs' state is created on the group's stream
                    with torch.cuda.stream(stream):
                        jobs = [Job(spec) for spec in chunk]
                        grp = JobGroup(jobs, stream)
                    grp.capture()
                    groups.append(grp)
                    n_groups += 1
                    if len(jobs) > 1:
                        shared += len(jobs)
                    n_active += len(jobs)
                    for j in jobs:
                        j.fp = footprint_bytes(j.spec)
                        if on_admit:
                            on_admit(j)
            max_conc = max(max_conc, n_active)
            if not groups:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival * time_scale - (time.perf_counter() - t0))))
                continue
            progressed = False
            still = []
            for grp in groups:
                if not grp.poll():
                    still.append(grp)
                    continue
                if grp.done:
                    progressed = True
                    fin = time.perf_counter() - t0
                    step_lat.extend(grp.latencies_ms)
                    for j in grp.jobs:
                        s = j.spec
                        finished_at[s.job_id] = fin - s.arrival * time_scale
                        total_samples += j.steps_done * s.n_samples
                        total_steps += j.steps_done
                        e_fin = float(np.mean([st.mean for st in j.stats[-10:]]))
                        finals[s.job_id] = e_fin
                        if s.dim == 1:
                            e0 = tfim_chain_exact_energy(s.L, s.h, s.J)
                            accuracy.append({"job_id": s.job_id, "L": s.L, "h": s.h, "alpha": s.alpha,
                                             "n_samples": s.n_samples, "energy": e_fin, "exact_e0": e0,
                                             "rel_err": abs(e_fin - e0) / abs(e0)})
                        if self.keep_params:
                            params[s.job_id] = j.vs.parameters_flat.cpu().numpy()
                        used -= j.fp
                        n_active -= 1
                        j.release()
                    continue
                grp.launch()
                progressed = True
                still.append(grp)
            groups = still
            if not progressed:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        jl = np.array(list(finished_at.values()))
        sl = np.array(step_lat)
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else float("nan")
        return ServingReport(
            n_jobs=len(specs), wall_s=wall, samples_per_s=total_samples / wall if wall > 0 else 0.0,
            steps_per_s=total_steps / wall if wall > 0 else 0.0,
            step_latency_ms_p50=pct(sl, 50), step_latency_ms_p99=pct(sl, 99),
            job_latency_s_p50=pct(jl, 50), job_latency_s_p99=pct(jl, 99),
            max_concurrent=max_conc, n_groups=n_groups, shared_launch_jobs=shared, final_energies=finals,
            final_params=params, rejected=rejected, accuracy=accuracy)
