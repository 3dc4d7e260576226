"""Pins of the oracle's energy gradient and of the SR / minSR right-hand side.

The reference checks the gradient of expect_and_grad against central
differences of the exact energy <psi|H|psi>/<psi|psi>
(test/driver/test_vmc.py:144-164 with central_diff_grad of
test/variational/finite_diff.py:32).  Here the same is done for the oracle:

* ``forces_to_grad`` evaluated with the exact weights |psi(s)|^2/Z over the
  full basis equals the finite-difference gradient of the energy from the
  dense Kronecker-product Hamiltonian (which shares no code with the
  connected-element generators);
* O_L^T dv of ``ngd.prepare`` with the same weights equals that gradient, so
  the scaling of dv (2 (E - <E>) / sqrt(Ns) for samples) is pinned;
* the SR and minSR updates tend to grad / eps as eps grows;
* the uniform-weight sample estimate (what the GPU tests compare against)
  converges to the exact gradient within six standard errors (the
  reference's bound, 6/sqrt(Ns), for unit variance);
* a gradient with a dropped factor or without centring fails these checks.
"""

import numpy as np
import pytest

from oracle import ffnn as of
from oracle import ngd as ongd
from oracle import operators as O
from oracle import rbm as orbm

SZSZ = np.diag([1.0, -1.0, -1.0, 1.0])
EXCH = np.array([[0, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 0]], dtype=np.float64)


def _chain(L):
    return np.array([(i, (i + 1) % L) for i in range(L)])


def _hamiltonian(kind, N):
    """(dense H from Kronecker products, conn function of the oracle)."""
    edges = _chain(N)
    if kind == "ising":
        h, J = 0.9, 1.0
        return O.dense_ising(N, edges, h, J), (lambda x: O.ising_conn(x, edges, h, J))
    # Heisenberg bonds J (sz sz + 2 exchange) (netket/operator/_heisenberg.py:97-124,
    # no sign rule), local index order of netket/operator/_local_operator/jax.py:37-55
    ops = [SZSZ + 2.0 * EXCH for _ in edges]
    acting = [tuple(int(v) for v in e) for e in edges]
    terms = O.canonical_terms(ops, acting)
    return O.dense_from_terms(N, ops, acting), (lambda x: O.local_conn(x, terms))


def _model(kind, N, seed):
    rng = np.random.default_rng(seed)
    if kind == "rbm":
        M = N
        flat = rng.normal(scale=0.25, size=M + N * M + N)
        make = lambda f: orbm.RBMParams(f, N, M)
        return flat, make, orbm.log_psi, orbm.jacobian, orbm.local_energy
    H1, H2 = 6, 5
    flat = rng.normal(scale=0.35, size=H1 + N * H1 + H2 + H1 * H2)
    make = lambda f: of.FFNNParams(f, N, H1, H2)
    return flat, make, of.log_psi, of.jacobian, of.local_energy


def _exact_energy(flat, make, log_psi, states, idx, Hd):
    lp = log_psi(make(flat), states)
    psi = np.zeros(Hd.shape[0])
    psi[idx] = np.exp(lp - lp.max())
    return float(psi @ (Hd @ psi) / (psi @ psi))


def _central_diff_grad(flat, make, log_psi, states, idx, Hd, eps=1e-5):
    g = np.empty_like(flat)
    for k in range(flat.size):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        g[k] = (_exact_energy(fp, make, log_psi, states, idx, Hd)
                - _exact_energy(fm, make, log_psi, states, idx, Hd)) / (2 * eps)
    return g


CASES = [("rbm", "ising"), ("rbm", "heisenberg"), ("ffnn", "ising"), ("ffnn", "heisenberg")]


def _setup(model_kind, ham_kind, N=8, seed=11):
    flat, make, log_psi, jac, eloc_fn = _model(model_kind, N, seed)
    Hd, conn = _hamiltonian(ham_kind, N)
    states = O.states_all(N)
    idx = O.state_index(states)
    p = make(flat)
    lp = log_psi(p, states)
    w = np.exp(2 * (lp - lp.max()))
    w /= w.sum()
    Ojac = jac(p, states)
    e = eloc_fn(p, states, conn)
    g_fd = _central_diff_grad(flat, make, log_psi, states, idx, Hd)
    return dict(flat=flat, p=p, w=w, O=Ojac, e=e, g_fd=g_fd, states=states, conn=conn,
                eloc_fn=eloc_fn, jac=jac, Hd=Hd, idx=idx, log_psi=log_psi, make=make)


@pytest.fixture(scope="module", params=CASES, ids=[f"{a}-{b}" for a, b in CASES])
def case(request):
    return _setup(*request.param)


def test_exact_energy_is_weighted_local_energy(case):
    # <E>_w with the exact weights is <psi|H|psi>/<psi|psi> (E_loc pinned against H psi / psi)
    e_exact = _exact_energy(case["flat"], case["make"], case["log_psi"], case["states"], case["idx"],
                            case["Hd"])
    assert np.sum(case["w"] * case["e"]) == pytest.approx(e_exact, rel=1e-12, abs=1e-12)


def test_gradient_formula_equals_central_differences(case):
    g, _ = orbm.forces_to_grad(case["O"], case["e"], case["w"])
    scale = np.max(np.abs(case["g_fd"]))
    assert scale > 1e-2  # a non-trivial gradient
    np.testing.assert_allclose(g, case["g_fd"], rtol=1e-6, atol=1e-8 * scale)


def test_dv_scaling_gives_the_gradient(case):
    OL, dv = ongd.prepare(case["O"], case["e"], case["w"])
    np.testing.assert_allclose(OL.T @ dv, case["g_fd"], rtol=1e-6, atol=1e-8 * np.max(np.abs(case["g_fd"])))
    # O_L^T O_L is the covariance of the log-derivatives under the weights
    w, Oj = case["w"], case["O"]
    mu = w @ Oj
    cov = (Oj - mu).T @ (w[:, None] * (Oj - mu))
    np.testing.assert_allclose(OL.T @ OL, cov, rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("eps", [1e6, 1e8])
def test_sr_and_minsr_updates_tend_to_gradient_over_eps(case, eps):
    OL, dv = ongd.prepare(case["O"], case["e"], case["w"])
    g = case["g_fd"]
    x_sr, _, _ = ongd.sr_update(OL, dv, eps)
    x_srt, _, _ = ongd.srt_update(OL, dv, eps)
    bound = 10.0 * np.linalg.norm(OL, 2) ** 2 / eps + 1e-6
    for x in (x_sr, x_srt):
        assert np.linalg.norm(eps * x - g) <= bound * np.linalg.norm(g)


def test_sample_estimate_converges_to_exact_gradient(case):
    # test/driver/test_vmc.py:144-164: n_samples = 5e5, six standard errors, rel 1e-3
    rng = np.random.default_rng(3)
    ns = 500_000
    pick = rng.choice(case["w"].size, size=ns, p=case["w"])
    cnt = np.bincount(pick, minlength=case["w"].size)
    # the uniform estimate over the drawn samples, evaluated with the per-state
    # counts (identical to repeating each drawn state cnt times)
    g_mc, _ = orbm.forces_to_grad(case["O"], case["e"], cnt / ns)
    # six standard errors of the per-sample terms 2 (O_s - <O>)(E_s - <E>)
    # under |psi|^2 (the reference's 6/sqrt(Ns) assumes unit variance)
    w, Oj, e = case["w"], case["O"], case["e"]
    y = 2.0 * (Oj - w @ Oj) * (e - w @ e)[:, None]
    sd = np.sqrt(np.maximum(w @ y**2 - (w @ y) ** 2, 0.0))
    assert np.all(np.abs(g_mc - case["g_fd"]) <= 6 * sd / np.sqrt(ns) + 1e-3 * np.abs(case["g_fd"]))
    # and literally on repeated rows for a smaller draw
    small = pick[:20_000]
    g_rows, _ = orbm.forces_to_grad(Oj[small], e[small])
    assert np.all(np.abs(g_rows - case["g_fd"]) <= 6 * sd / np.sqrt(small.size) + 1e-3 * np.abs(case["g_fd"]))
    # a factor 2 too small would be far outside the bound
    assert np.any(np.abs(0.5 * g_mc - case["g_fd"]) > 6 * sd / np.sqrt(ns) + 1e-3 * np.abs(case["g_fd"]))


def test_expect_and_grad_uses_the_pinned_formula():
    # the sample path of the oracle (what the GPU tests compare with) is the
    # weighted formula at w = 1/Ns: repeat every basis state in proportion to
    # integer weights and compare with the weighted formula
    c = _setup("rbm", "ising", N=6, seed=2)
    rng = np.random.default_rng(9)
    reps = rng.integers(1, 5, size=c["w"].size)
    x = np.repeat(c["states"], reps, axis=0)
    x = x[: (x.shape[0] // 4) * 4].reshape(4, -1, 6)
    n = x.shape[0] * x.shape[1]
    kept = np.repeat(np.arange(c["w"].size), reps)[:n]
    _, g, e = orbm.expect_and_grad(c["p"], x, c["conn"])
    g_w, _ = orbm.forces_to_grad(c["O"], c["e"], np.bincount(kept, minlength=c["w"].size) / n)
    np.testing.assert_allclose(g, g_w, rtol=1e-12, atol=1e-13)


def test_planted_mistakes_are_caught(case):
    # plausible mistakes in the gradient formula fail the finite-difference pin
    w, Oj, e, g_fd = case["w"], case["O"], case["e"], case["g_fd"]
    tol = 1e-6 * np.max(np.abs(g_fd))
    no_factor2 = Oj.T @ (w * (e - w @ e))
    no_centring = 2.0 * Oj.T @ (w * e)
    OL, dv = ongd.prepare(Oj, e, w)
    dv_sqrt_missing = OL.T @ (dv / np.sqrt(w))  # dv scaled by 1/Ns instead of 1/sqrt(Ns)
    for bad in (no_factor2, no_centring, dv_sqrt_missing):
        assert np.max(np.abs(bad - g_fd)) > 100 * tol
