"""Serving loop: continuous admission, overlapping jobs, and a job's result is
the same alone and in the mix."""

import numpy as np
import pytest

from netket_amd import serving

pytestmark = pytest.mark.gpu


def _specs():
    specs = serving.make_stream(n_jobs=6, rate=1000.0, seed=3, n_steps=6)
    for s in specs:
        s.n_samples = 1024
        s.L = 16 if s.dim == 1 else 4
    return specs


@pytest.mark.parametrize("use_graph", [True, False])
def test_job_same_alone_and_in_mix(gpu, use_graph):
    specs = _specs()
    mix = serving.Server(hbm_budget_gb=50, n_streams=4, use_graph=use_graph,
                         keep_params=True).run(specs, time_scale=0.0)
    assert mix.max_concurrent > 1
    assert len(mix.final_params) == len(specs)
    target = specs[2]
    alone = serving.Server(hbm_budget_gb=50, n_streams=1, use_graph=use_graph,
                           keep_params=True).run([target], time_scale=0.0)
    assert all(np.isfinite(e) for e in mix.final_energies.values())
    assert np.all(np.isfinite(alone.final_params[target.job_id]))
    np.testing.assert_array_equal(alone.final_params[target.job_id], mix.final_params[target.job_id])
    assert alone.final_energies[target.job_id] == mix.final_energies[target.job_id]


def test_admission_respects_memory_budget(gpu):
    specs = _specs()
    budget_gb = 1.2 * max(serving.footprint_bytes(s) for s in specs) / 2**30
    rep = serving.Server(hbm_budget_gb=budget_gb, n_streams=4).run(specs, time_scale=0.0)
    # only one job fits at a time, yet all complete
    assert rep.max_concurrent == 1
    assert len(rep.final_energies) == len(specs)


def test_many_captured_jobs_keep_their_scratch(gpu):
    # more captured jobs than the stream handles torch hands out (32), with
    # sample counts growing along the stream, so later jobs need larger
    # scratch buffers than earlier ones captured on the same stream handle;
    # every job's graph must keep replaying into its own buffers
    specs = serving.make_stream(n_jobs=40, rate=1000.0, seed=5, n_steps=3)
    for k, s in enumerate(specs):
        s.dim, s.L, s.alpha = 1, 16, 1
        s.n_samples = 1024 * (1 + k // 10)
    mix = serving.Server(hbm_budget_gb=50, n_streams=4, use_graph=True, max_active=64,
                         keep_params=True).run(specs, time_scale=0.0)
    assert len(mix.final_params) == len(specs)
    assert all(np.isfinite(e) for e in mix.final_energies.values())
    for target in (specs[0], specs[1]):
        alone = serving.Server(hbm_budget_gb=50, n_streams=1, use_graph=True,
                               keep_params=True).run([target], time_scale=0.0)
        np.testing.assert_array_equal(alone.final_params[target.job_id], mix.final_params[target.job_id])
