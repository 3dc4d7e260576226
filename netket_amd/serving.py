"""Continuous batching of independent ground-state VMC jobs on one GPU.

NetKet runs one driver loop per process (netket/_src/driver/abstract_variational_driver.py:349-499);
this module runs many such loops on one device.  A request stream (Poisson
arrivals of transverse-field Ising jobs of uneven size) is admitted by HBM
footprint; every admitted job keeps its parameters, chains and solver
buffers resident on the GPU, and its SR step is captured once into a HIP
graph.  Each scheduling round replays one step of every active job, each on
one of a pool of HIP streams, so the work of different jobs overlaps on the
device; jobs that arrive meanwhile are admitted as soon as memory frees up.
A job's result depends only on its own seeds: the Philox counters of its
sampler live in its own device state, so it is the same alone and in the mix.
"""

from __future__ import annotations

import math
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
from .driver import VMC_SR
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
    the centred Jacobian (bf16 hi/lo), the FP64 S or T matrix and workspaces,
    doubled for the graph pool kept next to the eager warm-up buffers."""
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
    b += rows * ld * 4                    # X hi + lo
    b += n * n * 8                        # S or T
    b += (Ns // 256 + 1) * (Np + 1) * 8   # contraction partials
    b += n * 16 + (n // 64 + 1) * 64 * 64 * 8
    return int(2 * b) + (32 << 20)


class Job:
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
        self.steps_done = 0
        self.energies = []
        self.step_events = []
        self.captured = False

    def step(self, use_graph: bool = True):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        if use_graph and not self.captured:
            # two eager warm-up steps, then the capture (the captured step runs on replay)
            s.record()
            self.drv.capture_step(warmup=2)
            self.steps_done += 2
            self.captured = True
            e.record()
            self.step_events.append((s, e, 2))
            return
        s.record()
        st = self.drv.replay_step() if use_graph else self.drv._eager_step()
        e.record()
        self.step_events.append((s, e, 1))
        self.energies.append(st)
        self.steps_done += 1

    @property
    def done(self) -> bool:
        return self.steps_done >= self.spec.n_steps

    def release(self):
        self.drv = None
        self.vs = None


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
    final_energies: dict = field(default_factory=dict)
    final_params: dict = field(default_factory=dict)
    rejected: list = field(default_factory=list)


class Server:
    def __init__(self, hbm_budget_gb: float = 200.0, n_streams: int = 8, use_graph: bool = True,
                 max_active: int = 64, keep_params: bool = False):
        self.keep_params = keep_params
        self.budget = int(hbm_budget_gb * (1 << 30))
        self.n_streams = n_streams
        self.streams = None
        self.use_graph = use_graph
        self.max_active = max_active

    def run(self, specs, time_scale: float = 1.0, on_admit=None, on_progress=None,
            progress_every_s: float = 10.0) -> ServingReport:
        # a job larger than the whole budget can never be admitted: it is
        # reported as rejected instead of blocking the queue
        rejected = [s.job_id for s in specs if footprint_bytes(s) > self.budget]
        pending = sorted((s for s in specs if footprint_bytes(s) <= self.budget), key=lambda s: s.arrival)
        if pending and self.streams is None:
            self.streams = [torch.cuda.Stream() for _ in range(self.n_streams)]
        active = []
        used = 0
        t0 = time.perf_counter()
        admitted_at, finished_at = {}, {}
        step_lat = []
        total_samples = 0
        total_steps = 0
        max_conc = 0
        finals = {}
        params = {}
        k = 0
        last_report = time.perf_counter()
        while pending or active:
            if on_progress is not None and time.perf_counter() - last_report >= progress_every_s:
                last_report = time.perf_counter()
                on_progress(len(finished_at), len(active), len(pending))
            now = time.perf_counter() - t0
            while (pending and pending[0].arrival * time_scale <= now and len(active) < self.max_active
                   and used + footprint_bytes(pending[0]) <= self.budget):
                spec = pending.pop(0)
                stream = self.streams[k % len(self.streams)]
                k += 1
                # the job's state is created on its own stream, which is not
                # ordered with the default stream
                with torch.cuda.stream(stream):
                    job = Job(spec)
                job.stream = stream
                job.fp = footprint_bytes(spec)
                used += job.fp
                active.append(job)
                admitted_at[spec.job_id] = now
                if on_admit:
                    on_admit(job)
            max_conc = max(max_conc, len(active))
            if not active:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival * time_scale - (time.perf_counter() - t0)))
                continue
            # one round: one step of every active job, each on its own stream
            round_start = time.perf_counter()
            for job in active:
                with torch.cuda.stream(job.stream):
                    job.step(self.use_graph)
            for s in self.streams:
                s.synchronize()
            round_ms = (time.perf_counter() - round_start) * 1e3
            still = []
            for job in active:
                step_lat.append(round_ms)
                if job.done:
                    fin = time.perf_counter() - t0
                    finished_at[job.spec.job_id] = fin - job.spec.arrival * time_scale
                    total_samples += job.steps_done * job.spec.n_samples
                    total_steps += job.steps_done
                    finals[job.spec.job_id] = float(np.mean([e.mean for e in job.energies[-10:]]))
                    if self.keep_params:
                        params[job.spec.job_id] = job.vs.parameters_flat.cpu().numpy()
                    used -= job.fp
                    job.release()
                else:
                    still.append(job)
            active = still
        wall = time.perf_counter() - t0
        jl = np.array(list(finished_at.values()))
        sl = np.array(step_lat)
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else float("nan")
        return ServingReport(
            n_jobs=len(specs), wall_s=wall, samples_per_s=total_samples / wall if wall > 0 else 0.0,
            steps_per_s=total_steps / wall if wall > 0 else 0.0,
            step_latency_ms_p50=pct(sl, 50), step_latency_ms_p99=pct(sl, 99),
            job_latency_s_p50=pct(jl, 50), job_latency_s_p99=pct(jl, 99),
            max_concurrent=max_conc, final_energies=finals, final_params=params, rejected=rejected)
