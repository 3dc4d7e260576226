"""Admission of the serving loop on the host (no GPU): a job whose HBM
footprint exceeds the whole budget is rejected instead of blocking the queue."""

from netket_amd import serving


def test_oversized_jobs_are_rejected_not_waited_for():
    specs = serving.make_stream(n_jobs=5, rate=100.0, seed=1, n_steps=3)
    need = min(serving.footprint_bytes(s) for s in specs)
    srv = serving.Server(hbm_budget_gb=(need - 1) / (1 << 30))
    rep = srv.run(specs)
    assert sorted(rep.rejected) == [s.job_id for s in specs]
    assert rep.final_energies == {} and srv.streams is None


def test_footprint_grows_with_job_size():
    small = serving.JobSpec(job_id=0, arrival=0.0, dim=1, L=16, h=1.0, alpha=1, n_samples=1024)
    big = serving.JobSpec(job_id=1, arrival=0.0, dim=2, L=12, h=1.0, alpha=4, n_samples=8192)
    assert serving.footprint_bytes(big) > 10 * serving.footprint_bytes(small)


def test_tfim_chain_exact_energy_equals_exact_diagonalisation():
    # the closed form the serving report compares each 1-D job with: even and
    # odd (frustrated, J > 0) periodic chains, against Lanczos on the oracle's
    # restated operator
    import numpy as np
    import pytest

    import netket_amd as nk
    from oracle import operators as O

    for L in (5, 6, 7, 9, 10):
        for h in (0.5, 1.0, 1.7):
            g = nk.graph.Chain(L)
            hi = nk.hilbert.Spin(0.5, N=L)
            H = nk.operator.Ising(hi, g, h=h, J=1.0)
            st = O.states_all(L)
            Hs = O.sparse_from_conn(st, lambda x: O.ising_conn(x, H.edges, h, 1.0))
            assert serving.tfim_chain_exact_energy(L, h) == pytest.approx(O.lanczos(Hs)[0], abs=1e-9)
    # the configured L = 16, h = 1 value
    assert serving.tfim_chain_exact_energy(16, 1.0) == pytest.approx(-20.404594474756657, abs=1e-12)
    assert np.isfinite(serving.tfim_chain_exact_energy(127, 0.8))
