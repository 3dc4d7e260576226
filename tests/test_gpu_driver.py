"""End-to-end VMC on the GPU: convergence to the exact ground state and the
SR == minSR equivalence of whole runs (test/driver/test_srt.py:98-211)."""

import numpy as np
import pytest

import netket_amd as nk
from netket_amd.optimizer import solver as gsolver
from oracle import operators as O

pytestmark = pytest.mark.gpu


def _tfim_state(L, n_chains, n_samples, n_discard, seed, alpha=1, dtype=np.float64):
    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L)
    H = nk.operator.Ising(hi, g, h=1.0, J=1.0)
    sa = nk.sampler.MetropolisLocal(hi, n_chains=n_chains)
    vs = nk.vqs.MCState(sa, nk.models.RBM(alpha=alpha, param_dtype=dtype), n_samples=n_samples,
                        n_discard_per_chain=n_discard, seed=seed)
    return H, vs


def test_vmc_sr_tfim_l16_converges(gpu):
    # The TFIM L=16 workload as configured: PBC h=J=1, FP64 RBM alpha=1,
    # MetropolisLocal 16 chains, 1008 samples, 16 discarded sweeps, SR
    # (QGTJacobianDense, diag_shift 0.01, Cholesky), SGD lr 0.01, seed 0.  The
    # run is deterministic, so the measured errors of the 100-step window mean
    # are asserted: about 2.6e-4 after the configured 300 steps, and within
    # 1e-4 of the exact energy by step 1000 (first at step 576), 6.0e-5 at 3000.
    H, vs = _tfim_state(16, 16, 1008, 16, seed=0)
    e0 = O.tfim_chain_e0(16, 1.0)
    sr = nk.optimizer.SR(nk.optimizer.QGTJacobianDense, solver=gsolver.cholesky, diag_shift=0.01)
    drv = nk.VMC(H, nk.optimizer.Sgd(0.01), variational_state=vs, preconditioner=sr)
    log = nk.logging.RuntimeLog()
    drv.run(3000, out=log)
    e = np.asarray(log.data["Energy"]["Mean"])
    rel = lambda k: abs(float(np.mean(e[k - 100:k])) - e0) / abs(e0)
    first = next(k for k in range(100, 3001) if rel(k) <= 1e-4)
    print(f"rel err: {rel(300):.2e} at 300, {rel(1000):.2e} at 1000, {rel(3000):.2e} at 3000; "
          f"first within 1e-4 at step {first}", flush=True)
    assert e[0] > e[-1]
    assert rel(300) < 5e-4
    assert first <= 1000
    assert rel(3000) <= 1e-4


def test_vmc_sr_and_minsr_runs_agree(gpu):
    # same initial state and sampler seed: SR and NTK updates give the same trajectory
    energies = {}
    for ntk in (False, True):
        H, vs = _tfim_state(8, 8, 64, 4, seed=3)
        drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.05), diag_shift=0.01, variational_state=vs,
                        use_ntk=ntk, linear_solver=gsolver.cholesky)
        log = nk.logging.RuntimeLog()
        drv.run(5, out=log)
        energies[ntk] = log.data["Energy"]["Mean"]
        params = vs.parameters_flat.cpu().numpy()
        energies[(ntk, "p")] = params
    np.testing.assert_allclose(energies[False], energies[True], rtol=0, atol=1e-10)
    np.testing.assert_allclose(energies[(False, "p")], energies[(True, "p")], rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("ntk", [False, True])
def test_graph_replay_matches_eager(gpu, ntk):
    # a captured step replays sampling (device-side Philox counter), E_loc,
    # Jacobian, Gram, Cholesky and update exactly as the eager path does
    import torch

    def make():
        H, vs = _tfim_state(10, 16, 256, 4, seed=11, dtype=np.float32)
        drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.05), diag_shift=0.01, variational_state=vs,
                        use_ntk=ntk, linear_solver=gsolver.cholesky)
        return drv, vs

    d1, v1 = make()
    e1 = [d1._eager_step().mean for _ in range(5)]
    d2, v2 = make()
    d2.capture_step(warmup=2)
    e2 = []
    for _ in range(3):
        e2.append(d2.replay_step().mean)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(v1.parameters_flat.cpu().numpy(), v2.parameters_flat.cpu().numpy())
    np.testing.assert_array_equal(np.array(e1[2:]), np.array(e2))


def test_mcstate_serialization_round_trip_and_json_log(gpu, tmp_path):
    # serialize_MCState / deserialize_MCState (netket/vqs/mc/mc_state/state.py:963-1016):
    # a restored state continues the chains exactly; JsonLog writes the .log history
    # and the variables as .mpack (netket/_src/logging/json_log.py:28-237)
    import json

    H, vs = _tfim_state(10, 8, 256, 2, seed=3, dtype=np.float32)
    drv = nk.VMC(H, nk.optimizer.Sgd(0.02), variational_state=vs,
                 preconditioner=nk.optimizer.SR(nk.optimizer.QGTJacobianDense, diag_shift=0.01))
    prefix = tmp_path / "run"
    drv.run(3, out=nk.logging.JsonLog(prefix, save_params_every=1, write_every=1))
    data = json.load(open(str(prefix) + ".log"))
    assert data["Energy"]["iters"] == [0, 1, 2] and len(data["Energy"]["Mean"]) == 3
    saved = nk.serialization.msgpack_restore(open(str(prefix) + ".mpack", "rb").read())
    for k, v in vs.parameters.items():
        if isinstance(v, dict):
            for kk, vv in v.items():
                np.testing.assert_array_equal(saved["params"][k][kk], vv.cpu().numpy())
        else:
            np.testing.assert_array_equal(saved["params"][k], v.cpu().numpy())

    blob = nk.serialization.to_bytes(vs)
    _, other = _tfim_state(10, 8, 128, 5, seed=99, dtype=np.float32)
    nk.serialization.from_bytes(other, blob)
    np.testing.assert_array_equal(other.parameters_flat.cpu().numpy(), vs.parameters_flat.cpu().numpy())
    assert other.n_samples == vs.n_samples and other.n_discard_per_chain == vs.n_discard_per_chain
    np.testing.assert_array_equal(other.sampler_state.sigma.cpu().numpy(), vs.sampler_state.sigma.cpu().numpy())
    s1 = vs.sample().cpu().numpy()
    s2 = other.sample().cpu().numpy()
    np.testing.assert_array_equal(s1, s2)


def test_chunk_size_local_energies_identical(gpu):
    # MCState.chunk_size (netket/vqs/mc/mc_state/state.py:471-512) bounds the
    # samples per launch; the local energies do not depend on it
    H, vs = _tfim_state(12, 16, 1024, 2, seed=4, dtype=np.float32)
    e_all = vs.local_estimators(H).cpu().numpy()
    with pytest.warns(UserWarning):
        vs.chunk_size = 100
    e_chunk = vs.local_estimators(H).cpu().numpy()
    np.testing.assert_array_equal(e_chunk, e_all)
    np.testing.assert_array_equal(vs.local_estimators(H, chunk_size=64).cpu().numpy(), e_all)
    for bad in (0, -4, 2.5):
        with pytest.raises(ValueError):
            vs.chunk_size = bad
    vs.chunk_size = None
    assert vs.chunk_size is None
