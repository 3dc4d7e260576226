"""Serving loop: continuous admission, shared launches for jobs of one shape,
groups that progress independently, and a job's result is the same alone and
in the mix."""

import ctypes as C

import numpy as np
import pytest
import torch

from netket_amd import _native as nat
from netket_amd import serving

pytestmark = pytest.mark.gpu


def _specs():
    # six jobs of distinct shapes
    specs = serving.make_stream(n_jobs=6, rate=1000.0, seed=3, n_steps=6)
    shapes = [(1, 16, 1), (1, 16, 2), (2, 4, 1), (2, 4, 2), (1, 20, 1), (1, 24, 4)]
    for s, (dim, L, alpha) in zip(specs, shapes):
        s.dim, s.L, s.alpha, s.n_samples = dim, L, alpha, 1024
    return specs


def _same_shape(specs, idx, dim=1, L=16, alpha=2, ns=1024):
    for k in idx:
        specs[k].dim, specs[k].L, specs[k].alpha, specs[k].n_samples = dim, L, alpha, ns


@pytest.mark.parametrize("shape", [(1, 16, 2, 1024), (2, 12, 4, 4096)])
def test_group_same_alone_and_in_mix(gpu, shape):
    # three jobs of one shape share their sweep (and, for the small model, local
    # energy) launches; each ends bit-identical to the job run on its own.  The
    # 12x12 alpha=4 model (W = 330 KB) takes the wide sampler and the per-job
    # tiled local energies.
    print(f"case serving group shape {shape}", flush=True)
    specs = _specs()
    _same_shape(specs, [1, 2, 4], *shape)
    for s in specs:
        s.n_steps = 5
    mix = serving.Server(hbm_budget_gb=80, n_streams=4, keep_params=True).run(specs, time_scale=0.0)
    assert mix.shared_launch_jobs == 3 and mix.n_groups == len(specs) - 2
    assert mix.max_concurrent == len(specs)
    for target in (specs[2], specs[0]):
        alone = serving.Server(hbm_budget_gb=80, n_streams=1, keep_params=True).run([target], time_scale=0.0)
        assert np.all(np.isfinite(alone.final_params[target.job_id]))
        np.testing.assert_array_equal(alone.final_params[target.job_id], mix.final_params[target.job_id])
        assert alone.final_energies[target.job_id] == mix.final_energies[target.job_id]


def test_batched_kernels_equal_single_launches(gpu):
    # the shared sweep and local-energy launches through the C ABI against one
    # launch per job: identical chains, acceptance counts and local energies
    import netket_amd as nk
    from netket_amd import _ops

    N, alpha, C_, n_jobs = 20, 2, 64, 3
    hi = nk.hilbert.Spin(0.5, N=N)
    g = nk.graph.Chain(N)
    model = nk.models.RBM(alpha=alpha, param_dtype=np.float32)
    rng = np.random.default_rng(0)
    flats = [torch.as_tensor(rng.normal(scale=0.2, size=model.layout(N)["n_params"]).astype(np.float32),
                             device="cuda") for _ in range(n_jobs)]
    ops = [nk.operator.Ising(hi, g, h=h, J=1.0) for h in (0.6, 1.0, 1.5)]
    sa = nk.sampler.MetropolisLocal(hi, n_chains=C_)
    res = {}
    for mode in ("single", "batched"):
        states = [sa.init_state(model, f, seed=10 + k) for k, f in enumerate(flats)]
        outs = []
        for st in states:
            st.log_prob = torch.empty(C_, dtype=torch.float32, device="cuda")
        samples = [torch.empty((C_, 3, N), dtype=torch.int8, device="cuda") for _ in range(n_jobs)]
        eloc = [torch.empty(C_ * 3, dtype=torch.float64, device="cuda") for _ in range(n_jobs)]
        if mode == "single":
            for k in range(n_jobs):
                s, states[k] = sa.sample_fused(model, flats[k], states[k], 2, 3)
                samples[k].copy_(s)
                eloc[k].copy_(nk._engine.local_energies(model, flats[k], ops[k], samples[k].reshape(-1, N)))
        else:
            jobs = (nat.NkaSweepJob * n_jobs)()
            ej = (nat.NkaElocJob * n_jobs)()
            keeps = []
            for k in range(n_jobs):
                st = states[k]
                jobs[k].model = model.native(flats[k], N)
                jobs[k].sigma, jobs[k].log_prob = st.sigma.data_ptr(), st.log_prob.data_ptr()
                jobs[k].samples, jobs[k].log_prob_samples = samples[k].data_ptr(), None
                jobs[k].n_accepted = st.n_accepted_proc.data_ptr()
                jobs[k].seed, jobs[k].counter, jobs[k].counter_dev = st.seed, st.counter, st.counter_dev.data_ptr()
                kind, pp, keep = _ops.op_struct(ops[k])
                keeps.append(keep)
                ej[k].model = model.native(flats[k], N)
                ej[k].op = nat.NkaIsing.from_buffer_copy(C.cast(pp, C.POINTER(nat.NkaIsing)).contents)
                ej[k].sigma, ej[k].eloc, ej[k].log_psi = samples[k].data_ptr(), eloc[k].data_ptr(), None
            jd = torch.frombuffer(bytearray(bytes(jobs)), dtype=torch.uint8).cuda()
            ed = torch.frombuffer(bytearray(bytes(ej)), dtype=torch.uint8).cuda()
            shape = model.native(flats[0], N)
            nat.call("nka_rbm_metropolis_batched", _ops.ptr(jd), n_jobs, C.byref(shape), nat.RULE_LOCAL, None, 0,
                     C_, sa.sweep_size, 2, 3, 2.0, _ops.stream())
            ising0 = nat.NkaIsing.from_buffer_copy(ej[0].op)
            assert nat.lib().nka_rbm_local_energy_batchable(C.byref(shape), C.byref(ising0)) == 1
            nat.call("nka_rbm_local_energy_batched", _ops.ptr(ed), n_jobs, C.byref(shape), C.byref(ising0),
                     C_ * 3, _ops.stream())
        torch.cuda.synchronize()
        res[mode] = ([s.cpu().numpy() for s in samples], [e.cpu().numpy() for e in eloc],
                     [st.n_accepted_proc.cpu().numpy() for st in states],
                     [int(st.counter_dev.cpu()[0]) for st in states])
    for a, b in zip(res["single"], res["batched"]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_admission_respects_memory_budget(gpu):
    # jobs of one footprint and a budget for one of them: one job at a time,
    # yet all complete
    specs = _specs()
    _same_shape(specs, range(len(specs)), 1, 24, 4, 1024)
    budget_gb = 1.2 * serving.footprint_bytes(specs[0]) / 2**30
    rep = serving.Server(hbm_budget_gb=budget_gb, n_streams=4).run(specs, time_scale=0.0)
    assert rep.max_concurrent == 1 and rep.n_groups == len(specs)
    assert len(rep.final_energies) == len(specs)


def test_many_captured_groups_keep_their_scratch(gpu):
    # more captured groups than the stream handles torch hands out (32), with
    # sample counts growing along the stream, so later jobs need larger
    # scratch buffers than earlier ones captured on the same stream handle;
    # every group's graph must keep replaying into its own buffers
    specs = serving.make_stream(n_jobs=40, rate=1000.0, seed=5, n_steps=3)
    for k, s in enumerate(specs):
        s.dim, s.L, s.alpha = 1, 16, 1
        s.n_samples = 1024 * (1 + k // 10)
    mix = serving.Server(hbm_budget_gb=50, n_streams=4, max_active=64, keep_params=True,
                         batch=False).run(specs, time_scale=0.0)
    assert len(mix.final_params) == len(specs) and mix.n_groups == len(specs)
    assert all(np.isfinite(e) for e in mix.final_energies.values())
    for target in (specs[0], specs[1]):
        alone = serving.Server(hbm_budget_gb=50, n_streams=1, keep_params=True).run([target], time_scale=0.0)
        np.testing.assert_array_equal(alone.final_params[target.job_id], mix.final_params[target.job_id])


def test_report_latencies_and_accuracy(gpu):
    # step latencies come from the events around each step; every 1-D job is
    # compared with the closed-form energy of its chain
    specs = _specs()
    for s in specs:
        s.n_steps = 30
    rep = serving.Server(hbm_budget_gb=50, n_streams=4).run(specs, time_scale=0.0)
    assert 0.0 < rep.step_latency_ms_p50 <= rep.step_latency_ms_p99
    one_d = [s for s in specs if s.dim == 1]
    assert len(rep.accuracy) == len(one_d)
    for a in rep.accuracy:
        assert a["exact_e0"] == pytest.approx(serving.tfim_chain_exact_energy(a["L"], a["h"]))
        assert a["energy"] >= a["exact_e0"] - 0.05 * abs(a["exact_e0"])  # variational, up to noise
