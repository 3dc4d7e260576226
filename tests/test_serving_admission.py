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
