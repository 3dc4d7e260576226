#!/usr/bin/env python
"""Benchmark of the VMC step on one MI355X.

Workload (timed): one VMC_SR step of the 1-D antiferromagnetic Heisenberg
chain, L=64, PBC, J=1, total Sz=0, Marshall sign rule; FP32 RBM alpha=2
(8384 parameters); MetropolisExchange on nearest-neighbour bonds (64
clusters), 4096 chains, 16384 samples per step, 5 discarded sweeps (NetKet's
default), sweep size 64; SR with the dense S (8384 x 8384) formed from the
RBM's spin-pair x activation-pair products on bf16 MFMA (FP64 centring) and
an FP64 Cholesky solve; diag_shift 0.01, SGD lr 0.01.
Synthetic data: parameters drawn N(0, 0.01^2) from seed 0, chains from the
sampler.  A step is sample -> E_loc -> statistics -> Jacobian -> S -> solve
-> update, all on the GPU.

Reported beside it, each outside the timed region:
* per-stage GPU times from HIP events on the launch stream, and the roofline
  of the largest kernel;
* the plain FP64 GPU step (the reference algorithm in PyTorch FP64,
  bench_plain_fp64.py) on the same workload: ``vs_baseline`` is the speed-up
  over it;
* accuracy: the TFIM L=16 workload (FP64) against its exact energy, run by
  both implementations for the same number of steps; the L=64 workload's
  energy after 500 steps against the Bethe-ansatz ground-state energy;
* the CPU restatement of the reference algorithm (FP64 NumPy/LAPACK) on one
  step of the same workload; its local energies, S and solution on the GPU
  step's own samples give the FP64-oracle error of each GPU result;
* the J1-J2 10x10 network minSR step, the TFIM 20x20 alpha=8 minSR step and
  the 256-job serving stream, each timed in this same run.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6
# Ground-state energy of the 64-site periodic Heisenberg chain (Pauli matrices,
# J = 1) from the Bethe equations; the same computation reproduces exact
# diagonalisation for L <= 14 and the tutorial's L = 22 value to 1e-12.
HEIS_L64_E0 = -113.49720554141149


def build_heisenberg_l64(seed=0):
    import numpy as np

    import netket_amd as nk
    from netket_amd.optimizer import solver

    L = 64
    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L, total_sz=0)
    H = nk.operator.Heisenberg(hi, g)
    sa = nk.sampler.MetropolisExchange(hi, graph=g, n_chains=4096)
    model = nk.models.RBM(alpha=2, param_dtype=np.float32)
    vs = nk.vqs.MCState(sa, model, n_samples=16384, n_discard_per_chain=5, seed=seed)
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=0.01, variational_state=vs,
                    use_ntk=False, linear_solver=solver.cholesky)
    return drv, H, vs


def build_j1j2_ffnn(seed=0, n_chains=8192):
    """J1-J2 Heisenberg on the 10x10 periodic square lattice (J1=1, J2=0.5,
    Sz=0, no sign rule), FP32 FFNN 100->400->400 (log_cosh, sum pooled,
    N(0, 0.01^2) init), MetropolisExchange over the 200 NN + 200 NNN bonds,
    16384 samples from 8192 chains (2 per chain, 5 discarded sweeps of 100),
    minSR with the structured NTK (16384^2 on bf16x3 tiles, FP64 output),
    diag_shift 1e-4, Cholesky."""
    import numpy as np

    import netket_amd as nk
    from netket_amd.optimizer import solver

    g = nk.graph.Square(10, max_neighbor_order=2)
    hi = nk.hilbert.Spin(0.5, N=g.n_nodes, total_sz=0)
    H = nk.operator.Heisenberg(hi, g, J=[1.0, 0.5], sign_rule=[False, False])
    sa = nk.sampler.MetropolisExchange(hi, graph=g, n_chains=n_chains)
    model = nk.models.FFNN(hidden_dims=(400, 400), param_dtype=np.float32)
    vs = nk.vqs.MCState(sa, model, n_samples=16384, n_discard_per_chain=5, seed=seed)
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=1e-4, variational_state=vs, use_ntk=True,
                    linear_solver=solver.cholesky)
    return drv, H, vs


def build_tfim2d_rbm(seed=0):
    """2-D TFIM on the 20x20 periodic square lattice, J=1, h=3.04, FP32 RBM
    alpha=8 (3200 hidden units, 1,283,600 parameters), MetropolisLocal 8192
    chains, 32768 samples (4 per chain, 5 discarded sweeps of 400), minSR with
    the structured NTK (32768^2 on bf16x3 tiles, FP64 output), diag_shift 1e-4,
    Cholesky."""
    import numpy as np

    import netket_amd as nk
    from netket_amd.optimizer import solver

    g = nk.graph.Square(20)
    hi = nk.hilbert.Spin(0.5, N=g.n_nodes)
    H = nk.operator.Ising(hi, g, h=3.04, J=1.0)
    sa = nk.sampler.MetropolisLocal(hi, n_chains=8192)
    model = nk.models.RBM(alpha=8, param_dtype=np.float32)
    vs = nk.vqs.MCState(sa, model, n_samples=32768, n_discard_per_chain=5, seed=seed)
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=1e-4, variational_state=vs, use_ntk=True,
                    linear_solver=solver.cholesky)
    return drv, H, vs


def run_workload(name, steps, warmup, norms=False):
    """Time one of the larger workloads (eager steps); returns its record.  With
    ``norms`` the energy, gradient and update norms of every step are added."""
    import torch

    from netket_amd import _ops

    builders = {"j1j2-ffnn": (build_j1j2_ffnn, 16384,
                              "j1j2-10x10 fp32-ffnn-100-400-400 metropolis-exchange-nn+nnn 8192ch "
                              "16384smp minsr-structured-ntk-16384 fp64-cholesky"),
                "tfim2d-rbm-a8": (build_tfim2d_rbm, 32768,
                                  "tfim-20x20-h3.04 fp32-rbm-a8 metropolis-local 8192ch 32768smp "
                                  "minsr-structured-ntk-32768 fp64-cholesky")}
    build, ns, desc = builders[name]
    drv, H, vs = build()
    for _ in range(warmup):
        step(drv)
    torch.cuda.synchronize()
    _ops.TIMER.reset()
    _ops.TIMER.enabled = True
    t0 = time.perf_counter()
    if norms:
        stats, grad_n, upd_n = [], [], []
        for _ in range(steps):
            drv.record = {}
            stats.append(step(drv))
            rec = drv.record
            # |grad| = |O_L^T dv| (the SR right-hand side without the shift), |update|
            grad_n.append(rec["rhs"].norm())
            upd_n.append(rec["dp"].norm())
        drv.record = None
    else:
        stats = [step(drv) for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _ops.TIMER.enabled = False
    stages = _ops.TIMER.summary()
    out = {
        "metric": "SR steps/s & samples/s",
        "value": steps / dt,
        "unit": "SR steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": dt / steps * 1e3,
        "samples_per_s": steps * ns / dt,
        "higher_is_better": True,
        "dtype": "f32 model, bf16x3 NTK (FP32 accumulate, FP64 epilogue), FP64 Cholesky",
        "data": "synthetic (random-init model, sampler-generated configurations)",
        "config": {"workload": desc, "n_samples": ns, "n_params": vs.n_parameters},
        "stages_ms_per_step": {k: round(v["total_ms"] / steps, 3) for k, v in sorted(
            stages.items(), key=lambda kv: -kv[1]["total_ms"])},
        "energy_per_step": [s.mean for s in stats],
        "acceptance": vs.sampler_state.acceptance,
    }
    if norms:
        out["dv_norm_per_step"] = [float(v) for v in grad_n]
        out["update_norm_per_step"] = [float(v) for v in upd_n]
    del drv, vs
    _ops.WORKSPACE.clear()
    torch.cuda.empty_cache()
    return out


def run_serving(n_jobs, n_steps, rate=4.0, verbose=True):
    """The serving workload: TFIM jobs of uneven size arriving as a Poisson
    stream (rate jobs/s), each running n_steps SR / minSR steps, admitted by
    HBM footprint and stepped as captured HIP graphs on per-job streams."""
    from netket_amd import serving

    specs = serving.make_stream(n_jobs=n_jobs, rate=rate, seed=0, n_steps=n_steps)
    def progress(done, active, pending):
        if verbose:
            print(f"serving: {done} done, {active} active, {pending} pending", file=sys.stderr, flush=True)

    rep = serving.Server(hbm_budget_gb=200.0, n_streams=8, use_graph=True, max_active=64).run(
        specs, on_progress=progress)
    out = {
        "metric": "p50/p99 job latency & aggregate samples/s (serving)",
        "value": rep.samples_per_s,
        "unit": "samples/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "job_latency_s_p50": rep.job_latency_s_p50,
        "job_latency_s_p99": rep.job_latency_s_p99,
        "step_latency_ms_p50": rep.step_latency_ms_p50,
        "step_latency_ms_p99": rep.step_latency_ms_p99,
        "steps_per_s": rep.steps_per_s,
        "wall_s": rep.wall_s,
        "max_concurrent": rep.max_concurrent,
        "groups": rep.n_groups,
        "jobs_in_shared_launches": rep.shared_launch_jobs,
        "accuracy_1d_jobs": _accuracy_summary(rep.accuracy),
        "dtype": "f32 model, bf16x3 / FP64 NTK, FP64 Cholesky",
        "data": "synthetic (random-init RBMs, sampler-generated configurations)",
        "config": {"workload": f"serving tfim-1d/2d fp32-rbm-a{{1,2,4}} {n_jobs} jobs x {n_steps} steps, "
                               f"poisson {rate}/s arrivals, samples/step in {{1024,4096,8192}}",
                   "n_jobs": n_jobs, "n_steps": n_steps, "rate": rate},
    }
    return out


def _accuracy_summary(acc):
    """Relative error of each 1-D job's final energy (mean of its last 10
    steps) against the exact energy of its chain: quantiles and counts."""
    import numpy as np

    if not acc:
        return None
    r = np.array([a["rel_err"] for a in acc])
    return {"n_jobs": int(r.size), "rel_err_p50": float(np.median(r)), "rel_err_p90": float(np.percentile(r, 90)),
            "rel_err_max": float(r.max()), "within_1e-2": int((r <= 1e-2).sum()),
            "within_1e-3": int((r <= 1e-3).sum())}


def step(drv):
    drv.reset_step()
    st, dp = drv.compute_loss_and_update()
    drv.update_parameters(dp)
    drv._step_count += 1
    return st


def accuracy_run(n_steps=3000, n_config=300, window=100):
    """The TFIM L=16 workload (FP64 RBM alpha=1, 16 chains, 1008 samples, 16 discarded
    sweeps, SR diag_shift 0.01 with Cholesky, SGD lr 0.01, seed 0) against the exact
    energy, as captured steps: the relative error of the mean energy over the last
    `window` steps at the configured n_config steps and at n_steps, and the first
    step at which the window mean is within 1e-4 of the exact energy."""
    import numpy as np
    import torch

    import netket_amd as nk
    from netket_amd.optimizer import solver

    L = 16
    e0 = -float(np.sum(np.sqrt(2.0 - 2.0 * np.cos(np.pi * (2 * np.arange(L) + 1) / L))))
    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L)
    H = nk.operator.Ising(hi, g, h=1.0, J=1.0)
    sa = nk.sampler.MetropolisLocal(hi, n_chains=16)
    vs = nk.vqs.MCState(sa, nk.models.RBM(alpha=1, param_dtype=np.float64), n_samples=1008,
                        n_discard_per_chain=16, seed=0)
    sr = nk.optimizer.SR(nk.optimizer.QGTJacobianDense, solver=solver.cholesky, diag_shift=0.01)
    lr = 0.01
    drv = nk.VMC(H, nk.optimizer.Sgd(lr), variational_state=vs, preconditioner=sr)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    drv.capture_step(warmup=2)
    stats = list(drv.warmup_stats)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(n_steps - len(stats)):
        stats.append(drv.replay_step())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    e = np.array([s.mean for s in stats])
    rel = lambda k: abs(float(np.mean(e[k - window:k])) - e0) / abs(e0)
    reached = next((k for k in range(window, n_steps + 1) if rel(k) <= 1e-4), None)
    sps = (n_steps - 2) / dt
    return {
        "workload": "tfim-1d-L16-h1 fp64-rbm-a1 metropolis-local 16ch 1008smp sr-cholesky "
                    f"eps=0.01 sgd-lr={lr}",
        "exact_e0": e0,
        f"rel_err_at_{n_config}_steps": rel(n_config),
        f"rel_err_at_{n_steps}_steps": rel(n_steps),
        "rel_err": rel(n_steps),
        "window": window,
        "first_step_within_1e-4": reached,
        "steps_per_s": sps,
        "time_to_1e-4_s": (reached / sps) if reached else None,
        "captured_steps": True,
        "setup_and_capture_s": t1 - t0,
    }


def heisenberg_l64_energy(n_steps=500, window=50):
    """The timed workload run for its configured 500 steps (captured steps, seed 0):
    the mean energy of the last `window` steps against the Bethe-ansatz E0."""
    import numpy as np
    import torch

    drv, H, vs = build_heisenberg_l64(seed=0)
    drv.capture_step(warmup=2)
    stats = list(drv.warmup_stats)
    for _ in range(n_steps - len(stats)):
        stats.append(drv.replay_step())
    torch.cuda.synchronize()
    e = np.array([s.mean for s in stats])
    tail = float(np.mean(e[-window:]))
    err = float(np.std(e[-window:]) / math.sqrt(window))
    return {"workload": "heisenberg-1d-L64-sz0 (the timed workload), seed 0, captured steps",
            "steps": n_steps, "exact_e0_bethe": HEIS_L64_E0, f"energy_mean_last_{window}": tail,
            "spread_of_mean": err, "rel_err": abs(tail - HEIS_L64_E0) / abs(HEIS_L64_E0),
            "energy_every_50_steps": [float(v) for v in e[::50]] + [float(e[-1])],
            "acceptance": vs.sampler_state.acceptance}


def plain_fp64_heisenberg_l64(steps=3, warmup=1):
    """SR steps/s of the plain FP64 GPU implementation on the timed workload."""
    import torch

    from bench_plain_fp64 import PlainRBMVMC

    v = PlainRBMVMC(64, 2, 4096, 16384, 5, ("heisenberg", 1.0), 0.01, 0.01, seed=0)
    for _ in range(warmup):
        v.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    es = [v.step() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"value": steps / dt, "unit": "SR steps/s", "ms_per_step": dt / steps * 1e3, "steps": steps,
           "energy_last": float(es[-1])}
    del v
    torch.cuda.empty_cache()
    return out


def plain_fp64_tfim16(max_steps=3000, budget_s=60.0, window=100):
    """The plain FP64 GPU implementation on the TFIM L=16 workload: steps/s, and
    the first step whose window mean is within 1e-4 of the exact energy (run
    until then, up to max_steps or budget_s)."""
    import numpy as np
    import torch

    from bench_plain_fp64 import PlainRBMVMC

    L = 16
    e0 = -float(np.sum(np.sqrt(2.0 - 2.0 * np.cos(np.pi * (2 * np.arange(L) + 1) / L))))
    v = PlainRBMVMC(L, 1, 16, 1008, 16, ("ising", 1.0, 1.0), 0.01, 0.01, seed=0)
    es = [v.step()]  # includes the capture of the sampler
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reached = None
    k = 1
    while k < max_steps and time.perf_counter() - t0 < budget_s:
        chunk = [v.step() for _ in range(100)]
        es.extend(chunk)
        k += len(chunk)
        e = torch.stack(es).cpu().numpy()
        for j in range(max(window, k - 100), k + 1):
            if abs(float(np.mean(e[j - window:j])) - e0) / abs(e0) <= 1e-4:
                reached = j
                break
        if reached:
            break
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    e = torch.stack(es).cpu().numpy()
    sps = (k - 1) / dt
    del v
    torch.cuda.empty_cache()
    return {"steps_run": k, "steps_per_s": sps, "first_step_within_1e-4": reached,
            "rel_err_last_window": abs(float(np.mean(e[-window:])) - e0) / abs(e0),
            "time_to_1e-4_s": (reached / sps) if reached else None}


def record_step(drv):
    """One eager step's inputs and results, copied to the host: the samples,
    parameters, local energies, S (lower triangle), F and the solution."""
    import numpy as np

    drv.record = {}
    drv.reset_step()
    st, dp = drv.compute_loss_and_update()
    rec = drv.record
    drv.record = None
    S = rec["system"].lower()
    out = {"samples": drv.state.samples.cpu().numpy().reshape(-1, drv.state.hilbert.size),
           "params": drv.state.parameters_flat.cpu().numpy().astype(np.float64),
           "eloc": rec["eloc"].cpu().numpy(), "S": np.tril(S.cpu().numpy()),
           "F": rec["rhs"].cpu().numpy(), "x": rec["solution"].cpu().numpy()}
    del S
    drv.update_parameters(dp)
    return out


def cpu_baseline_heisenberg_l64(rec, sample_budget_s: float = 30.0):
    """One CPU step of the reference algorithm (FP64) on the timed workload: the
    sampler on its own 4096 chains (5 + 4 sweeps of 64 full-forward exchange
    proposals), then local energies, Jacobian, S and the Cholesky solve on the
    GPU step's samples and parameters.  The stage results on those samples are
    the FP64 oracle of the GPU step's local energies, S and solution."""
    import numpy as np

    import netket_amd as nk
    from oracle import cpu_baseline as cb
    from oracle.rbm import RBMParams

    L, M = 64, 128
    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L, total_sz=0)
    H = nk.operator.Heisenberg(hi, g)
    terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
    clusters = nk.sampler.compute_clusters(g, 1)[0].astype(np.int64)
    rng = np.random.default_rng(0)
    C = 4096
    sigma = np.ones((C, L), dtype=np.int8)
    for c in range(C):
        sigma[c, rng.permutation(L)[: L // 2]] = -1
    timings = {}
    flat = rec["params"]
    t0 = time.perf_counter()
    cb.exchange_sweeps(RBMParams(flat, L, M), sigma, clusters, 5 + 4, L, rng, 5)
    timings["sample"] = time.perf_counter() - t0
    e, S, F, dp = cb.sr_stages(flat, L, M, rec["samples"].astype(np.int8), terms, 0.01, timings)
    step_s = sum(timings.values())
    Sg, xg, eg = rec["S"], rec["x"], rec["eloc"]
    lo = np.tril_indices_from(S)
    d = np.sqrt(np.diag(S) - 0.01)
    dS = np.abs(Sg[lo] - S[lo])
    errors = {
        "eloc_max_abs_err_rel": float(np.abs(eg - e).max() / np.abs(e).max()),
        "energy_mean_rel_err": float(abs(eg.mean() - e.mean()) / abs(e.mean())),
        "S_max_abs_err_rel": float(dS.max() / np.abs(S[lo]).max()),
        "S_max_err_over_sqrt_SiiSjj": float((dS / (np.outer(d, d)[lo] + 1e-300)).max()),
        "F_rel_err": float(np.linalg.norm(rec["F"] - F) / np.linalg.norm(F)),
        "x_rel_err": float(np.linalg.norm(xg - dp) / np.linalg.norm(dp)),
        "reference": "FP64 NumPy/LAPACK restatement of the reference algorithm on the GPU step's "
                     "own samples and (FP32) parameters",
    }
    return 1.0 / step_s, timings, errors


def roofline_for(stage: str, mean_ms: float):
    Ns, Np, n_chains = 16384, 8384, 4096
    secs = mean_ms * 1e-3
    if stage == "qgt_structured":
        # C = P Q^T: 2081 spin-pair rows x 8385 activation-pair rows over the
        # 16384 samples; each product runs twice on the matrix cores (Q = hi + lo)
        N, M = 64, 128
        na, nb = 1 + N + N * (N - 1) // 2, 1 + M + M * (M + 1) // 2
        flops = 2.0 * na * nb * Ns
        ach = flops / secs / 1e12
        return {"bound": "mfma", "achieved": ach, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / BF16_PEAK_TFLOPS, "traffic": None, "kernel": "nka_rbm_qgt (all launches)",
                "per_launch": f"2*{na}*{nb}*Ns = {flops:.3e} logical flop (x2 MFMA work for the bf16 hi + lo "
                              f"split of Q); the Jacobian Gram would take 2*Ns*Np^2 = {2.0 * Ns * Np * Np:.3e}",
                "mean_ms": mean_ms}
    if stage == "potrf":
        flops = Np**3 / 3.0
        ach = flops / secs / 1e12
        return {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "nka_potrf_f64 (all launches)",
                "per_launch": f"n^3/3 = {flops:.3e} flop, n = {Np}"}
    if stage == "metropolis":
        u = unit_utilisation("rbm_metropolis_group_kernel")
        props = n_chains * (5 + 4) * 64
        return {"bound": "valu", "achieved": u.get("valu_busy") if u else None, "peak": 1.0,
                "unit": "fraction of SIMD cycles issuing VALU (counters)",
                "frac": u.get("valu_busy") if u else None, "traffic": measured_traffic("rbm_metropolis_group_kernel"),
                "kernel": "rbm_metropolis_group_kernel",
                "per_launch": f"{props} proposals; W (32 KB) is staged in LDS, so HBM carries only the chains "
                              "and samples: the sweep is bound by the VALU work of the log-cosh updates and the "
                              "latency of the sequential proposals, not by HBM",
                "mean_ms": mean_ms}
    if stage == "local_energy":
        u = unit_utilisation("rbm_eloc_kernel")
        return {"bound": "salu", "achieved": u.get("salu_issue") if u else None, "peak": 1.0,
                "unit": "fraction of scalar-unit cycles issuing SALU (counters)",
                "frac": u.get("salu_issue") if u else None, "valu_busy": u.get("valu_busy") if u else None,
                "traffic": measured_traffic("rbm_eloc_kernel"), "kernel": "rbm_eloc_kernel",
                "per_launch": "16384 samples x 64 bonds; the walk over the operator tables (one scalar lookup chain "
                              "per bond) keeps the scalar unit the busiest, W is in LDS",
                "mean_ms": mean_ms}
    return None


def unit_utilisation(kernel: str):
    """MFMA / VALU / SALU busy fractions and the L2 hit rate of ``kernel`` from the
    committed counter summary (profiles/unit_utilisation.json), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "unit_utilisation.json")
    try:
        with open(path) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    for name, rec in kernels.items():
        if name.split("<")[0] == kernel:
            return rec
    return None


def copy_calibration(nbytes: int = 1 << 30, reps: int = 3):
    """A plain device-to-device copy of nbytes (reads nbytes, writes nbytes),
    the reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}))


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy calibration), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
        kernels = data["kernels"]
        if kernel in kernels:
            return kernels[kernel]["hbm_bytes_per_launch"]
        # template instances are listed as name<args>
        for name, rec in kernels.items():
            if name.split("<")[0] == kernel:
                return rec["hbm_bytes_per_launch"]
        return None
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic_per_call(kernels, per_call_kernel: str):
    """HBM bytes per call of an entry point that launches several kernels: the
    sum over ``kernels`` of bytes per launch x launches, divided by the number
    of calls, which is the dispatch count of ``per_call_kernel`` (launched once
    per call), from profiles/hbm_traffic.json; None when a kernel is missing."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    calls = sum(r["dispatches"] for n, r in recs.items() if n.split("<")[0] == per_call_kernel)
    if not calls:
        return None
    total = 0.0
    for k in kernels:
        found = [r for n, r in recs.items() if n.split("<")[0] == k]
        if not found:
            return None
        total += sum(r["hbm_bytes_per_launch"] * r["dispatches"] for r in found)
    return int(total / calls)


POTRF_KERNELS = ("panel_step_kernel", "gemm_nt_f64_big_kernel", "panel_factor_kernel", "gemm_big_reduce_kernel",
                 "trtri_blocks_kernel", "potf2_kernel", "panel_update_kernel")
QGT_KERNELS = ("qgt_spins_t_kernel", "qgt_spin_pairs_kernel", "qgt_hidden_means_kernel", "qgt_hidden_pairs_kernel",
               "gemm_exact_a_kernel", "qgt_unpack_kernel")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def guarded(name, fn, *a, **kw):
    """Run a reporting leg; a failure is reported in the JSON line (and its
    traceback on stderr) instead of losing the timed result."""
    import traceback

    try:
        return fn(*a, **kw)
    except Exception as e:  # noqa: BLE001 - reported, not hidden
        traceback.print_exc()
        log(f"{name} failed: {e!r}")
        return {"error": f"{type(e).__name__}: {e}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU baseline")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the J1-J2 FFNN, TFIM 20x20 and serving workloads")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream only")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured-graph replays")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--workload", default="heisenberg-l64",
                    choices=["heisenberg-l64", "j1j2-ffnn", "tfim2d-rbm-a8", "serving"],
                    help="the default Heisenberg L=64 SR step (with every report), or only one of "
                         "the larger workloads, printed as its own line")
    ap.add_argument("--norms", action="store_true",
                    help="with --workload j1j2-ffnn / tfim2d-rbm-a8: per-step gradient and update norms")
    ap.add_argument("--serving-jobs", type=int, default=256)
    ap.add_argument("--serving-steps", type=int, default=200)
    ap.add_argument("--workload-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the VMC step here runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from netket_amd import _ops

    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "serving":
        print(json.dumps(run_serving(args.serving_jobs, args.serving_steps)))
        return
    if args.workload != "heisenberg-l64":
        print(json.dumps(run_workload(args.workload, args.steps, args.warmup, norms=args.norms)))
        return
    log("timed workload: heisenberg L=64")
    drv, H, vs = build_heisenberg_l64()
    for _ in range(args.warmup):
        step(drv)
    torch.cuda.synchronize()
    _ops.TIMER.reset()
    _ops.TIMER.enabled = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [step(drv) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _ops.TIMER.enabled = False
    stages = _ops.TIMER.summary()
    dt_eager = t1 - t0
    # the same step captured once into a HIP graph and replayed
    dt_graph = float("inf")
    if not args.no_graph:
        drv.capture_step(warmup=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats_g = [drv.replay_step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt_graph = time.perf_counter() - t0
        stats = stats + stats_g
    dt = min(dt_eager, dt_graph)
    sps = args.steps / dt
    gpu_ms_total = sum(v["total_ms"] for v in stages.values()) / args.steps
    energies = [s.mean for s in stats]
    dom = max(stages.items(), key=lambda kv: kv[1]["total_ms"])
    # roofline of the dominant stage (the Cholesky factorisation of S); the
    # other stages are listed beside it
    roof = roofline_for("potrf", stages["potrf"]["mean_ms"])
    roof["traffic"] = measured_traffic_per_call(POTRF_KERNELS, "potf2_kernel")
    others = {k: roofline_for(k, stages[k]["mean_ms"]) for k in ("qgt_structured", "metropolis", "local_energy")
              if k in stages}
    if others.get("qgt_structured"):
        others["qgt_structured"]["traffic"] = measured_traffic_per_call(QGT_KERNELS, "qgt_spins_t_kernel")
    # one more eager step with its results kept for the oracle comparison
    rec = None if args.no_cpu_baseline else record_step(drv)
    del drv, vs
    _ops.WORKSPACE.clear()
    torch.cuda.empty_cache()

    acc = acc64 = None
    if not args.no_accuracy:
        log("accuracy: tfim L=16 (3000 captured steps)")
        acc = guarded("accuracy tfim16", accuracy_run)
        log("accuracy: heisenberg L=64 (500 captured steps) vs Bethe ansatz")
        acc64 = guarded("accuracy heisenberg64", heisenberg_l64_energy)
        _ops.WORKSPACE.clear()
        torch.cuda.empty_cache()
    plain = plain16 = None
    vs_baseline = None
    if not args.no_plain:
        log("plain FP64 GPU baseline: heisenberg L=64")
        plain = guarded("plain heisenberg64", plain_fp64_heisenberg_l64)
        if "value" in plain:
            vs_baseline = sps / plain["value"]
        if not args.no_accuracy:
            log("plain FP64 GPU baseline: tfim L=16 to 1e-4")
            plain16 = guarded("plain tfim16", plain_fp64_tfim16)
    cpu = errors = None
    if rec is not None:
        log("CPU restatement: one step on the GPU step's samples")
        cores = len(os.sched_getaffinity(0))
        res = guarded("cpu baseline", cpu_baseline_heisenberg_l64, rec)
        rate, tim, errors = res if isinstance(res, tuple) else (None, {}, res)
        threads = int(os.environ.get("OMP_NUM_THREADS", cores))
        cpu = None if rate is None else {"value": rate, "unit": "SR steps/s", "cores": min(cores, threads), "kind": "port",
               "sample": "1 full step of the same workload (L=64 Heisenberg, 4096 chains, 16384 samples, "
                         "S 8384^2) in FP64 NumPy/LAPACK with the reference algorithm: its sampler on its "
                         "own chains, the rest on the GPU step's samples; stage seconds "
                         f"{json.dumps({k: round(v, 2) for k, v in tim.items()})}"}
        del rec
    workloads = {}
    if not args.no_workloads:
        for name in ("j1j2-ffnn", "tfim2d-rbm-a8"):
            log(f"workload {name}")
            workloads[name] = guarded(name, run_workload, name, args.workload_steps, 2)
        if not args.no_serving:
            log(f"serving: {args.serving_jobs} jobs x {args.serving_steps} steps")
            workloads["serving"] = guarded("serving", run_serving, args.serving_jobs, args.serving_steps)
    out = {
        "metric": "SR steps/s & samples/s at <=1e-4 rel. energy err vs ED; p50/p99 job latency",
        "value": sps,
        "unit": "SR steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "ms_per_step_eager": dt_eager / args.steps * 1e3,
        "ms_per_step_graph": (dt_graph / args.steps * 1e3) if math.isfinite(dt_graph) else None,
        "samples_per_s": sps * 16384,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": vs_baseline,
        "vs_baseline_definition": "value / SR steps/s of the plain FP64 GPU step (the reference algorithm "
                                  "in PyTorch FP64 on this GPU: full forward per proposal, materialised "
                                  "connected elements, dense FP64 Jacobian, rocBLAS S, rocSOLVER Cholesky) "
                                  "on the same workload, measured in this run",
        "dtype": "f32 (S on bf16 MFMA: exact spin pairs x hi+lo activation pairs, FP32 accumulate; "
                 "FP64 centring, Cholesky, means)",
        "data": "synthetic (random-init RBM, sampler-generated configurations)",
        "config": {"workload": "heisenberg-1d-L64-sz0 fp32-rbm-a2 metropolis-exchange 4096ch "
                               "16384smp sr-dense-S8384 structured-bf16x2 fp64-cholesky",
                   "n_samples": 16384, "n_params": 8384, "n_chains": 4096, "n_discard": 5,
                   "diag_shift": 0.01, "lr": 0.01, "parallelism": "single-gpu"},
        "stages_ms_per_step": {k: round(v["total_ms"] / args.steps, 4) for k, v in sorted(
            stages.items(), key=lambda kv: -kv[1]["total_ms"])},
        "gpu_busy_ms_per_step": gpu_ms_total,
        "dominant_stage": dom[0],
        "roofline": roof,
        "roofline_other_stages": others,
        "cpu_baseline": cpu,
        "fp64_oracle_error": errors,
        "plain_fp64_gpu": {"heisenberg_l64": plain, "tfim_l16": plain16},
        "accuracy": acc,
        "accuracy_heisenberg_l64": acc64,
        "energy_first_last": [energies[0], energies[-1]],
        "workloads": workloads,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
