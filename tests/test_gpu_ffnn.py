"""GPU parity of the two-layer dense network path (J1-J2 workload ansatz)
against the oracle: log psi, local energies, the log-derivative contractions,
the structured NTK, the Jacobian, the SR / minSR update, Metropolis chains
and the sampled distribution."""

import numpy as np
import pytest
import torch

import netket_amd as nk
from netket_amd import _engine as eng
from netket_amd import _native as nat
from netket_amd.optimizer import solver as gsolver
from oracle import ffnn as of
from oracle import ngd as ongd
from oracle import operators as O
from oracle import sampler as osamp

pytestmark = pytest.mark.gpu


def _model(N, H1, H2, dtype, seed, scale, use_bias=True):
    model = nk.models.FFNN(hidden_dims=(H1, H2), param_dtype=dtype, use_bias=use_bias)
    n = model.layout(N)["n_params"]
    flat = np.random.default_rng(seed).normal(scale=scale, size=n).astype(dtype)
    return model, torch.as_tensor(flat, device="cuda"), of.FFNNParams(flat.astype(np.float64), N, H1, H2,
                                                                      use_bias)


def _rtol(dtype):
    return 1e-12 if dtype == np.float64 else 2e-5


@pytest.mark.parametrize("N,H1,H2,dtype", [(10, 12, 9, np.float64), (10, 12, 9, np.float32),
                                           (100, 400, 400, np.float32), (33, 70, 45, np.float64),
                                           (7, 1, 1, np.float64)])
def test_log_psi(gpu, N, H1, H2, dtype):
    print(f"case ffnn log_psi N={N} H=({H1},{H2}) {np.dtype(dtype)}", flush=True)
    model, flat, p = _model(N, H1, H2, dtype, N + H1, 0.3 if N < 50 else 0.05)
    x = np.random.default_rng(1).choice(np.array([-1, 1], dtype=np.int8), size=(77, N))
    lp = model.apply(flat, torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = of.log_psi(p, x)
    err = np.abs(lp - ref).max() / np.abs(ref).max()
    print(f"max rel err {err:.2e}", flush=True)
    assert err < _rtol(dtype)


def _j1j2(L):
    g = nk.graph.Square(L, max_neighbor_order=2)
    hi = nk.hilbert.Spin(0.5, N=g.n_nodes, total_sz=0)
    H = nk.operator.Heisenberg(hi, g, J=[1.0, 0.5], sign_rule=[False, False])
    terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
    return g, hi, H, (lambda x: O.local_conn(x, terms))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_local_energy_j1j2(gpu, dtype):
    print(f"case ffnn eloc J1-J2 4x4 {np.dtype(dtype)}", flush=True)
    g, hi, H, conn = _j1j2(4)
    N = hi.size
    model, flat, p = _model(N, 24, 20, dtype, 3, 0.2)
    x = np.stack([np.random.default_rng(s).permutation([1] * 8 + [-1] * 8) for s in range(300)]).astype(np.int8)
    e, lp = eng.local_energies(model, flat, H, torch.as_tensor(x, device="cuda"), log_psi=True)
    # the list of nonzero connected rows is built with atomics, so its order
    # varies; each row's value does not, and neither does the sum
    e2 = eng.local_energies(model, flat, H, torch.as_tensor(x, device="cuda"))
    np.testing.assert_array_equal(e2.cpu().numpy(), e.cpu().numpy())
    ref = of.local_energy(p, x, conn)
    tol = 1e-10 if dtype == np.float64 else 1e-4
    np.testing.assert_allclose(e.cpu().numpy(), ref, rtol=tol, atol=tol * np.abs(ref).max())
    np.testing.assert_allclose(lp.cpu().numpy(), of.log_psi(p, x), rtol=_rtol(dtype))


def test_local_energy_ising_small_workspace(gpu):
    # a workspace smaller than the whole batch takes several chunks
    print("case ffnn eloc ising chunks", flush=True)
    N = 12
    hi = nk.hilbert.Spin(0.5, N=N)
    H = nk.operator.Ising(hi, nk.graph.Chain(N), h=0.7, J=1.1)
    model, flat, p = _model(N, 16, 10, np.float64, 5, 0.3)
    x = np.random.default_rng(2).choice(np.array([-1, 1], dtype=np.int8), size=(200, N))
    m = model.native(flat, N)
    from netket_amd import _ops
    import ctypes as C
    kind, op_ptr, keep = _ops.op_struct(H)
    per = nat.lib().nka_ffnn_local_energy_workspace(C.byref(m), kind, op_ptr, 1)
    ws = torch.empty(int(per * 7 + 4096), dtype=torch.uint8, device="cuda")
    eloc = torch.empty(200, dtype=torch.float64, device="cuda")
    xt = torch.as_tensor(x, device="cuda")
    nat.call("nka_ffnn_local_energy", C.byref(m), kind, op_ptr, _ops.ptr(xt), 200, _ops.ptr(eloc), None,
             _ops.ptr(ws), ws.numel(), _ops.stream())
    ref = of.local_energy(p, x, lambda z: O.ising_conn(z, H.edges, 0.7, 1.1))
    np.testing.assert_allclose(eloc.cpu().numpy(), ref, rtol=1e-10, atol=1e-12)
    del keep


@pytest.mark.parametrize("dtype,use_bias", [(np.float64, True), (np.float64, False), (np.float32, True)])
def test_contract_jacobian_and_ntk(gpu, dtype, use_bias):
    print(f"case ffnn contract/jacobian/ntk {np.dtype(dtype)} bias={use_bias}", flush=True)
    N, H1, H2, ns = 11, 13, 9, 300
    model, flat, p = _model(N, H1, H2, dtype, 8, 0.3, use_bias)
    x = np.random.default_rng(9).choice(np.array([-1, 1], dtype=np.int8), size=(ns, N))
    sig = torch.as_tensor(x, device="cuda")
    f, _ = eng.prepare(model, flat, sig)
    O_ = of.jacobian(p, x)
    tol = 1e-11 if dtype == np.float64 else 3e-5
    # mean and a weighted, centred contraction
    ones = torch.full((1, ns), 1.0 / ns, dtype=torch.float64, device="cuda")
    mean = eng.contract(model, flat, sig, f, ones)[0]
    np.testing.assert_allclose(mean.cpu().numpy(), O_.mean(axis=0), rtol=tol, atol=tol)
    w = np.random.default_rng(10).normal(size=(2, ns))
    out = eng.contract(model, flat, sig, f, torch.as_tensor(w, device="cuda"), mean).cpu().numpy()
    Oc = O_ - O_.mean(axis=0)
    ref = w @ Oc
    np.testing.assert_allclose(out, ref, rtol=tol, atol=tol * np.abs(ref).max())
    # Jacobian, SP layout, FP64 output
    X = eng.jacobian(model, flat, sig, f, None, mean, 0.5, nat.LAYOUT_SP, nat.OUT_F64)
    np.testing.assert_allclose(X.hi[:, : p.n_params].cpu().numpy(), 0.5 * Oc, rtol=tol, atol=tol)
    # structured NTK vs the explicit product
    T = eng.ntk_structured(model, flat, sig, f, mean, 1.0 / ns, 1e-3, 0.1 / ns).cpu().numpy()
    Tref = Oc @ Oc.T / ns + 1e-3 * np.eye(ns) + 0.1 / ns
    lo = np.tril_indices(ns)
    du = np.sqrt(np.sum(O_ * O_, axis=1) / ns)
    dc = np.sqrt(np.sum(Oc * Oc, axis=1) / ns)
    bound = (1e-12 * np.outer(dc, dc) + 1e-14 * np.outer(du, du))[lo] + 1e-15
    if dtype == np.float32:
        bound = (3e-5 * np.outer(dc, dc) + 1e-6 * np.outer(du, du))[lo] + 1e-9
    err = np.abs(T[lo] - Tref[lo])
    print(f"ntk max err / bound {(err / bound).max():.2e}", flush=True)
    assert np.all(err <= bound)


@pytest.mark.parametrize("ntk_method", ["structured", "jacobian"])
def test_vmc_sr_update_matches_oracle(gpu, ntk_method):
    print(f"case ffnn sr/minsr {ntk_method}", flush=True)
    g, hi, H, conn = _j1j2(4)
    N = hi.size
    sa = nk.sampler.MetropolisExchange(hi, graph=g, n_chains=16)
    model = nk.models.FFNN(hidden_dims=(10, 8), param_dtype=np.float64)
    vs = nk.vqs.MCState(sa, model, n_samples=128, n_discard_per_chain=4, seed=3)
    flat_np = np.random.default_rng(4).normal(scale=0.2, size=vs.n_parameters)
    vs.parameters = torch.as_tensor(flat_np, device="cuda")
    p = of.FFNNParams(flat_np, N, 10, 8)
    x = vs.samples.cpu().numpy().reshape(-1, N)
    e = of.local_energy(p, x, conn)
    OL, dv = ongd.prepare(of.jacobian(p, x), e)
    ref, _, _ = ongd.sr_update(OL, dv, 1e-3)
    out = {}
    for ntk in (False, True):
        drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=1e-3, variational_state=vs, use_ntk=ntk,
                        linear_solver=gsolver.cholesky, ntk_method=ntk_method)
        st, dp = drv.compute_loss_and_update()
        out[ntk] = dp.cpu().numpy()
    scale = np.abs(ref).max()
    np.testing.assert_allclose(out[False], ref, rtol=0, atol=1e-8 * scale)
    np.testing.assert_allclose(out[True], ref, rtol=0, atol=1e-8 * scale)


def test_vmc_minsr_spring_momentum_two_steps(gpu):
    # minSR with momentum 0.8 through the network's structured Jacobian-vector
    # product, two chained steps against the oracle (netket/_src/ngd/srt.py:34-82)
    print("case ffnn minsr spring momentum", flush=True)
    g, hi, H, conn = _j1j2(4)
    N = hi.size
    sa = nk.sampler.MetropolisExchange(hi, graph=g, n_chains=16)
    model = nk.models.FFNN(hidden_dims=(10, 8), param_dtype=np.float64)
    vs = nk.vqs.MCState(sa, model, n_samples=64, n_discard_per_chain=4, seed=5)
    vs.parameters = torch.as_tensor(np.random.default_rng(6).normal(scale=0.2, size=vs.n_parameters),
                                    device="cuda")
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.05), diag_shift=1e-2, variational_state=vs, use_ntk=True,
                    momentum=0.8, linear_solver=gsolver.cholesky)
    old = np.zeros(vs.n_parameters)
    for step in range(2):
        drv.reset_step()
        p = of.FFNNParams(vs.parameters_flat.cpu().numpy(), N, 10, 8)
        x = vs.samples.cpu().numpy().reshape(-1, N)
        st, dp = drv.compute_loss_and_update()
        OL, dv = ongd.prepare(of.jacobian(p, x), of.local_energy(p, x, conn))
        ref, _, _ = ongd.srt_update(OL, dv, 1e-2, momentum=0.8, old_updates=old)
        np.testing.assert_allclose(dp.cpu().numpy(), ref, rtol=0, atol=1e-8 * np.abs(ref).max())
        old = ref
        drv.update_parameters(dp)
        drv._step_count += 1


@pytest.mark.parametrize("rule,H1,H2,scale", [("local", 9, 7, 0.4), ("exchange", 9, 7, 0.4),
                                               ("exchange", 420, 30, 0.05), ("exchange-weighted", 9, 7, 0.4)])
def test_chains_identical_to_oracle_fp64(gpu, rule, H1, H2, scale):
    # H1 = 420: 53 first-layer units per thread, the 56-unit sampler instance
    print(f"case ffnn chains {rule} H1={H1}", flush=True)
    N = 10
    model, flat, p = _model(N, H1, H2, np.float64, 3, scale)
    g = nk.graph.Chain(N)
    weights = None
    if rule == "local":
        hi = nk.hilbert.Spin(0.5, N=N)
        sa = nk.sampler.MetropolisLocal(hi, n_chains=40, sweep_size=7)
        clusters = None
    elif rule == "exchange":
        hi = nk.hilbert.Spin(0.5, N=N, total_sz=0)
        sa = nk.sampler.MetropolisExchange(hi, graph=g, n_chains=40, sweep_size=7)
        clusters = sa.rule.clusters
    else:
        hi = nk.hilbert.Spin(0.5, N=N, total_sz=0)
        sa = nk.sampler.MetropolisExchange(hi, graph=g, d_max=2, probabilities=[0.4, 1.0], n_chains=40,
                                           sweep_size=7)
        clusters, weights = sa.rule.clusters, sa.rule.weights
        rule = "exchange"
    st = sa.init_state(model, flat, seed=99)
    sigma0 = st.sigma.cpu().numpy().copy()
    counter0 = st.counter
    samples, st = sa.sample_fused(model, flat, st, 3, 9)
    ref, ref_sigma, ref_acc = osamp.metropolis(p, sigma0, rule=rule, clusters=clusters, n_discard=3,
                                               chain_length=9, sweep_size=7, seed=99, counter=counter0,
                                               log_psi_fn=of.log_psi, weights=weights)
    np.testing.assert_array_equal(samples.cpu().numpy(), ref)
    np.testing.assert_array_equal(st.sigma.cpu().numpy(), ref_sigma)
    np.testing.assert_array_equal(st.n_accepted_proc.cpu().numpy(), ref_acc)
    np.testing.assert_allclose(st.log_prob.cpu().numpy(), 2 * of.log_psi(p, ref_sigma), rtol=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_distribution_matches_exact(gpu, dtype):
    print(f"case ffnn distribution {np.dtype(dtype)}", flush=True)
    N = 8
    model, flat, p = _model(N, 6, 5, dtype, 7, 0.3)
    hi = nk.hilbert.Spin(0.5, N=N, total_sz=0)
    sa = nk.sampler.MetropolisExchange(hi, graph=nk.graph.Chain(N), n_chains=1024)
    states = O.states_all(N, n_down=4)
    st = sa.init_state(model, flat, seed=5)
    samples, st = sa.sample_fused(model, flat, st, 100, 200)
    s = samples.cpu().numpy().reshape(-1, N)
    keys = O.state_index(s)
    order = np.argsort(O.state_index(states))
    skeys = O.state_index(states)[order]
    pos = np.searchsorted(skeys, keys)
    assert np.all(skeys[pos] == keys), "sample outside the Hilbert space"
    counts = np.bincount(pos, minlength=len(states)).astype(float)
    lp = 2 * of.log_psi(p, states[order])
    probs = np.exp(lp - lp.max())
    probs /= probs.sum()
    q = counts / counts.sum()
    m = q > 0
    kl = float(np.sum(q[m] * np.log(q[m] / probs[m])))
    print(f"KL = {kl:.3e}", flush=True)
    assert kl < 4e-3


def test_vmc_sr_ffnn_tfim_converges(gpu):
    # TFIM chain L=8, h=1: minSR with the structured NTK reaches the exact energy
    L = 8
    hi = nk.hilbert.Spin(0.5, N=L)
    H = nk.operator.Ising(hi, nk.graph.Chain(L), h=1.0, J=1.0)
    sa = nk.sampler.MetropolisLocal(hi, n_chains=64)
    vs = nk.vqs.MCState(sa, nk.models.FFNN(hidden_dims=(16, 16), param_dtype=np.float64, kernel_init_std=0.1),
                        n_samples=1024, n_discard_per_chain=8, seed=1)
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.05), diag_shift=1e-3, variational_state=vs, use_ntk=True,
                    linear_solver=gsolver.cholesky)
    log = nk.logging.RuntimeLog()
    drv.run(150, out=log)
    e = np.asarray(log.data["Energy"]["Mean"])
    e0 = O.tfim_chain_e0(L, 1.0)
    tail = float(np.mean(e[-20:]))
    rel = abs(tail - e0) / abs(e0)
    print(f"E(last 20) = {tail:.6f} E0 = {e0:.6f} rel {rel:.2e}", flush=True)
    assert rel < 5e-3
