"""Pins of the CPU oracle against the reference's known answers and against
independent constructions (no GPU)."""

import numpy as np
import pytest

from oracle import ngd as ongd
from oracle import operators as O
from oracle import philox
from oracle import rbm as orbm
from oracle import stats as ostats


# --- Philox4x32-10: Random123 known-answer vectors ------------------------------
@pytest.mark.parametrize(
    "ctr,key,expect",
    [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ],
)
def test_philox_kat(ctr, key, expect):
    assert list(map(int, philox.philox4x32_10(ctr, key))) == expect


# --- Ising: exact energies of the reference ---------------------------------------
def _chain_edges(L):
    return np.array([(i, (i + 1) % L) for i in range(L)])


def test_tfim_closed_form_matches_reference_values():
    # netket/exact.py:61-63 and docs/tutorials/gs-ising.ipynb:282
    assert O.tfim_chain_e0(8, 1.0) == pytest.approx(-10.25166179, abs=5e-9)
    assert O.tfim_chain_e0(20, 1.0) == pytest.approx(-25.49098969, abs=5e-9)


def test_ising_conn_lanczos_L8_three_lowest():
    # netket/exact.py:61-63: lowest three eigenvalues of TFIM L=8, h=J=1
    L = 8
    st = O.states_all(L)
    H = O.sparse_from_conn(st, lambda x: O.ising_conn(x, _chain_edges(L), 1.0, 1.0))
    w = O.lanczos(H, k=3)
    np.testing.assert_allclose(w, [-10.25166179, -10.05467898, -8.69093921], atol=5e-8)


def test_ising_conn_lanczos_L16_closed_form():
    L = 16
    st = O.states_all(L)
    H = O.sparse_from_conn(st, lambda x: O.ising_conn(x, _chain_edges(L), 1.0, 1.0))
    assert O.lanczos(H)[0] == pytest.approx(-20.404594474756657, abs=1e-9)


def test_ising_conn_equals_kronecker_hamiltonian():
    L = 6
    rng = np.random.default_rng(0)
    edges = np.array([(0, 1), (1, 2), (2, 5), (3, 4), (4, 0), (2, 3)])
    h, J = 0.7, -1.3
    st = O.states_all(L)
    Hs = O.sparse_from_conn(st, lambda x: O.ising_conn(x, edges, h, J)).toarray()
    Hd = O.dense_ising(L, edges, h, J)
    # the conn diagonal is J * (integer sum), the Kronecker one a sum of +-J terms
    np.testing.assert_allclose(Hs, Hd, rtol=0, atol=1e-13)
    # row 0 of the connected elements is x itself, rows r>0 flip site r-1
    x = st[rng.integers(0, len(st), 5)]
    xp, mels = O.ising_conn(x, edges, h, J)
    np.testing.assert_array_equal(xp[:, 0], x)
    for r in range(1, L + 1):
        d = (xp[:, r] != x).sum(axis=1)
        assert np.all(d == 1) and np.all(xp[:, r, r - 1] == -x[:, r - 1])
    assert np.all(mels[:, 1:] == -h)


# --- Heisenberg / local operators --------------------------------------------------
def test_heisenberg_L22_sz0_reference_energy():
    # docs/tutorials/gs-heisenberg.ipynb:163: E0 = -39.147522607062776
    import netket_amd as nk

    L = 22
    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L, total_sz=0)
    H = nk.operator.Heisenberg(hi, g)
    terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
    st = O.states_all(L, n_down=11)
    Hs = O.sparse_from_conn(st, lambda x: O.local_conn(x, terms))
    assert O.lanczos(Hs)[0] == pytest.approx(-39.147522607062776, abs=1e-9)


@pytest.mark.parametrize("L", [4, 8, 10, 12, 14])
def test_bethe_ansatz_equals_exact_diagonalisation(L):
    # the Bethe-equation energy (oracle/bethe.py) against Lanczos on the
    # restated operator in the Sz = 0 sector, and the tutorial's L = 22 value
    import netket_amd as nk
    from oracle import bethe

    g = nk.graph.Chain(L)
    hi = nk.hilbert.Spin(0.5, N=L, total_sz=0)
    H = nk.operator.Heisenberg(hi, g)
    terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
    st = O.states_all(L, n_down=L // 2)
    Hs = O.sparse_from_conn(st, lambda x: O.local_conn(x, terms))
    assert bethe.heisenberg_e0(L) == pytest.approx(O.lanczos(Hs)[0], abs=1e-9)


def test_bethe_ansatz_l22_and_l64():
    from oracle import bethe

    assert bethe.heisenberg_e0(22) == pytest.approx(-39.147522607062776, abs=1e-10)
    # the L = 64 value the benchmark compares with; per site it approaches
    # 4 (1/4 - ln 2) from below as 1/L^2 (finite-size correction)
    e64 = bethe.heisenberg_e0(64)
    assert e64 == pytest.approx(-113.49720554141149, abs=1e-9)
    assert 4 * (0.25 - np.log(2)) - 0.01 < e64 / 64 < 4 * (0.25 - np.log(2))


def test_local_conn_equals_kronecker_elementwise():
    rng = np.random.default_rng(1)
    N = 6
    ops, aon = [], []
    for _ in range(7):
        k = int(rng.integers(1, 4))
        sites = list(rng.choice(N, size=k, replace=False))
        m = rng.normal(size=(2**k, 2**k))
        m[np.abs(m) < 0.6] = 0.0
        ops.append(m + m.T)
        aon.append(sites)
    terms = O.canonical_terms(ops, aon)
    st = O.states_all(N)
    Hs = O.sparse_from_conn(st, lambda x: O.local_conn(x, terms, constant=0.25)).toarray()
    Hd = O.dense_from_terms(N, ops, aon, constant=0.25)
    np.testing.assert_allclose(Hs, Hd, atol=1e-12)


def test_local_conn_drops_zero_diagonal_and_pads():
    # Heisenberg on a 4-chain: Neel-like states have zero diagonal, which is dropped
    import netket_amd as nk

    g = nk.graph.Chain(4)
    hi = nk.hilbert.Spin(0.5, N=4)
    H = nk.operator.Heisenberg(hi, g)
    terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
    x = np.array([[1, -1, 1, -1], [1, 1, -1, -1], [1, 1, 1, 1]], dtype=np.int8)
    xp, mels, n = O.local_conn(x, terms)
    assert xp.shape == (3, 5, 4)
    # state 0: all 4 bonds antiparallel: diag = -4 (kept), 4 exchanges
    assert n[0] == 5 and mels[0, 0] == -4.0
    # state 1: 2 parallel + 2 antiparallel: diag 0 dropped, 2 exchanges, padding (x, 0)
    assert n[1] == 2 and np.all(mels[1, 2:] == 0) and np.all(xp[1, 2:] == x[1])
    # state 2: all parallel: only the diagonal
    assert n[2] == 1 and mels[2, 0] == 4.0


# --- RBM ----------------------------------------------------------------------------
def _rbm(N, M, seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    flat = rng.normal(scale=scale, size=M + N * M + N)
    return orbm.RBMParams(flat, N, M)


def test_log_cosh_stable_form():
    x = np.array([-30.0, -1.5, -1e-3, 0.0, 2e-3, 0.7, 40.0])
    np.testing.assert_allclose(orbm.log_cosh(x), np.log(np.cosh(x)), rtol=1e-13, atol=1e-15)


def test_local_energy_equals_dense_H_psi():
    N, M = 6, 12
    p = _rbm(N, M)
    st = O.states_all(N)
    edges = _chain_edges(N)
    e = orbm.local_energy(p, st, lambda x: O.ising_conn(x, edges, 0.8, 1.0))
    Hd = O.dense_ising(N, edges, 0.8, 1.0)
    psi = np.exp(orbm.log_psi(p, st))
    order = np.argsort(O.state_index(st))
    psi_full = np.empty_like(psi)
    psi_full[O.state_index(st)] = psi
    expect = (Hd @ psi_full)[O.state_index(st)] / psi
    np.testing.assert_allclose(e, expect, rtol=1e-12)
    del order


def test_jacobian_finite_differences():
    N, M = 5, 7
    p = _rbm(N, M, seed=3)
    rng = np.random.default_rng(4)
    x = rng.choice([-1, 1], size=(4, N)).astype(np.int8)
    O_ = orbm.jacobian(p, x)
    flat = np.concatenate([p.b, p.W.ravel(), p.a])
    eps = 1e-6
    for k in range(flat.size):
        fp, fm = flat.copy(), flat.copy()
        fp[k] += eps
        fm[k] -= eps
        d = (orbm.log_psi(orbm.RBMParams(fp, N, M), x) - orbm.log_psi(orbm.RBMParams(fm, N, M), x)) / (2 * eps)
        np.testing.assert_allclose(O_[:, k], d, rtol=1e-6, atol=1e-8)


# --- statistics -----------------------------------------------------------------------
def test_stats_mean_var_and_error_branches():
    rng = np.random.default_rng(5)
    data = rng.normal(size=(64, 20))
    st = ostats.statistics(data)
    assert st["mean"] == pytest.approx(data.mean(), rel=1e-14)
    assert st["variance"] == pytest.approx(data.var(), rel=1e-14)
    # 64 chains >= 32: batch estimator, error = sqrt(var(chain means)/n_chains)
    assert st["error_of_mean"] == pytest.approx(np.sqrt(data.mean(axis=1).var() / 64), rel=1e-13)
    # few chains, long chains: block estimator (l_block = L // 32)
    d2 = rng.normal(size=(4, 256))
    st2 = ostats.statistics(d2)
    blocks = d2.reshape(-1, 8).mean(axis=1)
    assert st2["error_of_mean"] == pytest.approx(np.sqrt(blocks.var() / blocks.size), rel=1e-13)
    # few chains, short chains (4 x 4: 4 chain means, 16 blocks): no valid estimator
    assert np.isnan(ostats.statistics(rng.normal(size=(4, 4)))["error_of_mean"])
    # split R-hat: halves of each chain
    halves = data.reshape(128, 10).mean(axis=1).var()
    assert st["R_hat"] == pytest.approx(np.sqrt(19 / 20 + halves / data.var()), rel=1e-13)


# --- SR / minSR ------------------------------------------------------------------------
def test_sr_equals_minsr_woodbury():
    # test/driver/test_srt.py:98-211: SR and the NTK formulation give the same update
    rng = np.random.default_rng(6)
    O_ = rng.normal(size=(40, 25))
    e = rng.normal(size=40)
    OL, dv = ongd.prepare(O_, e)
    x1, _, _ = ongd.sr_update(OL, dv, 1e-3)
    x2, _, _ = ongd.srt_update(OL, dv, 1e-3)
    np.testing.assert_allclose(x1, x2, rtol=1e-9, atol=1e-12)


def test_sr_equals_minsr_with_spring_momentum():
    # with momentum mu the minSR update O_L^T (T + eps)^-1 (dv - mu O_L u) + mu u equals the
    # SR update (S + eps)^-1 (O_L^T dv + eps mu u) (Woodbury), over two chained steps
    rng = np.random.default_rng(16)
    u_sr = u_srt = np.zeros(25)
    for step in range(2):
        OL, dv = ongd.prepare(rng.normal(size=(40, 25)), rng.normal(size=40))
        x1, _, _ = ongd.sr_update(OL, dv, 1e-2, momentum=0.8, old_updates=u_sr)
        x2, _, _ = ongd.srt_update(OL, dv, 1e-2, momentum=0.8, old_updates=u_srt)
        np.testing.assert_allclose(x1, x2, rtol=1e-9, atol=1e-12)
        if step == 1:
            # the momentum term changes the update
            x0, _, _ = ongd.sr_update(OL, dv, 1e-2)
            assert np.linalg.norm(x1 - x0) > 1e-3 * np.linalg.norm(x0)
        u_sr, u_srt = x1, x2


def test_cholesky_and_cg_and_pinv():
    rng = np.random.default_rng(7)
    A = rng.normal(size=(30, 30))
    A = A @ A.T + 0.1 * np.eye(30)
    b = rng.normal(size=30)
    x = ongd.cholesky_solve(A, b)
    np.testing.assert_allclose(A @ x, b, rtol=1e-10, atol=1e-10)
    xc, k = ongd.cg(lambda v: A @ v, b, tol=1e-10)
    np.testing.assert_allclose(xc, x, rtol=1e-6, atol=1e-8)
    assert 0 < k <= 300
    # pinv_smooth on a singular matrix equals the least-squares minimum-norm solution
    U = rng.normal(size=(20, 5))
    S = U @ U.T
    bb = S @ rng.normal(size=20)
    xp = ongd.pinv_smooth(S, bb, rtol=1e-10, rtol_smooth=0.0)
    np.testing.assert_allclose(xp, np.linalg.lstsq(S, bb, rcond=1e-10)[0], rtol=1e-7, atol=1e-9)


# --- sampler: the proposal rules satisfy detailed balance w.r.t. |psi|^2 -----------------
def test_weighted_pick_gives_each_cluster_its_weight():
    # k uniform over [0, W): cluster i must be chosen by exactly w_i of the W draws
    from oracle import sampler as osamp

    w = np.array([3, 1, 0, 7, 2, 5], dtype=np.int64)
    picks = np.array([osamp.pick_weighted(w, k) for k in range(int(w.sum()))])
    np.testing.assert_array_equal(np.bincount(picks, minlength=w.size), w)


def test_exchange_rule_integer_weights_and_distance_probabilities():
    import netket_amd as nk

    g = nk.graph.Chain(6)
    r = nk.sampler.ExchangeRule(graph=g, d_max=2, probabilities=[1.0, 0.25])
    d = np.asarray(g.distances())[r.clusters[:, 0], r.clusters[:, 1]]
    np.testing.assert_array_equal(r.probabilities, np.where(d == 1, 1.0, 0.25))
    np.testing.assert_array_equal(r.weights, np.where(d == 1, 1 << 20, 1 << 18).astype(np.uint32))
    assert nk.sampler.integer_weights([1e-9, 1.0]).tolist() == [1, 1 << 20]
    with pytest.raises(ValueError):
        nk.sampler.ExchangeRule(graph=g, probabilities=[0.0])
    with pytest.raises(TypeError):
        nk.sampler.ExchangeRule(clusters=[(0, 1), (1, 2)], probabilities=[1.0])
    with pytest.raises(ValueError):
        nk.sampler.ExchangeRule(graph=g, d_max=2, probabilities=[1.0])



@pytest.mark.parametrize("rule", ["local", "exchange", "exchange-weighted"])
def test_metropolis_transition_matrix_is_stationary_on_psi2(rule):
    from oracle import sampler as osamp

    N = 8
    p = _rbm(N, 8, seed=7, scale=0.5)
    states = O.states_all(N, n_down=4 if rule != "local" else None)
    idx = {int(k): i for i, k in enumerate(O.state_index(states))}
    lp = 2 * orbm.log_psi(p, states)
    # nearest and next-nearest neighbours; weighted: distance 2 three times as likely
    clusters = np.sort(np.array([(i, (i + d) % N) for d in (1, 2) for i in range(N)]), axis=1)
    weights = np.where(np.arange(2 * N) < N, 1.0, 3.0) if rule == "exchange-weighted" else None
    T = np.zeros((len(states), len(states)))
    for i, x in enumerate(states):
        if rule == "local":
            moves, corrs, probs = osamp.local_moves(x)
        else:
            moves, corrs, probs = osamp.exchange_moves(x, clusters, weights)
        assert sum(probs) == pytest.approx(1.0)
        for xp, c, q in zip(moves, corrs, probs):
            j = idx[int(O.state_index(xp[None])[0])]
            T[i, j] += min(1.0, np.exp(lp[j] - lp[i] + c)) * q
        T[i, i] += 1.0 - T[i].sum()
    pi = orbm.exact_probability(p, states)
    np.testing.assert_allclose(pi @ T, pi, rtol=1e-10, atol=1e-14)
    # detailed balance
    F = pi[:, None] * T
    np.testing.assert_allclose(F, F.T, atol=1e-14)
