"""The decomposition behind the structured dense S of the RBM, restated in
NumPy FP64 on the CPU and checked against the explicit product of the
oracle's Jacobian (netket/_src/ngd/sr.py:50-55: O_L^T O_L).

Each RBM log-derivative is (augmented spin) x (augmented tanh): hidden bias
b_j = 1 * t_j, kernel W_ij = sigma_i * t_j, visible bias a_i = sigma_i * 1.
The uncentred sums are entries of C = P Q^T over the samples, P the distinct
spin pairs and Q the distinct pairs of centred activations t' = t - c, and
S = scale (U_pq - U_p U_q / ns) with

  U(u,u',w,w') = C[r,(w,w')] + c_w C[r,(-1,w')] + c_w' C[r,(w,-1)] + c_w c_w' C[r,(-1,-1)],
  U_p = U(u_p, -1, w_p, -1).

The GPU kernels (nka_rbm_qgt) compute exactly this with P and Q in bf16 and C
in FP32; here every step is FP64, so the two sides agree to rounding.  A wrong
pair index, a dropped centring term or a sign error fails it.
"""

import numpy as np
import pytest

from oracle import rbm as orbm


def _spin_row(u, v, N):
    if u == v:
        return 0
    if u < 0:
        return 1 + v
    if v < 0:
        return 1 + u
    i, k = min(u, v), max(u, v)
    return 1 + N + k * (k - 1) // 2 + i


def _hidden_row(w, x, M):
    if w < 0 and x < 0:
        return 0
    if w < 0:
        return 1 + x
    if x < 0:
        return 1 + w
    j, l = min(w, x), max(w, x)
    return 1 + M + l * (l + 1) // 2 + j


def _param(u, w, N, M, hb, vb):
    off_k = M if hb else 0
    if u < 0:
        return w if (hb and w >= 0) else -1
    if w < 0:
        return off_k + N * M + u if vb else -1
    return off_k + u * M + w


def structured_s(p: orbm.RBMParams, x: np.ndarray, scale: float) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    ns, N = x.shape
    M = p.M
    hb, vb = p.b is not None, p.a is not None
    t = np.tanh(orbm.theta(p, x))
    c = t.mean(axis=0)
    tp = t - c
    # P rows: 1, sigma_i, sigma_i sigma_k (i < k); Q rows: 1, t'_j, t'_j t'_l (j <= l)
    rp = 1 + N + N * (N - 1) // 2
    rq = 1 + M + M * (M + 1) // 2
    P = np.zeros((rp, ns))
    Q = np.zeros((rq, ns))
    P[0] = 1.0
    Q[0] = 1.0
    for i in range(N):
        P[1 + i] = x[:, i]
        for k in range(i + 1, N):
            P[_spin_row(i, k, N)] = x[:, i] * x[:, k]
    for j in range(M):
        Q[1 + j] = tp[:, j]
        for l in range(j, M):
            Q[_hidden_row(j, l, M)] = tp[:, j] * tp[:, l]
    C = P @ Q.T
    n_par = (M if hb else 0) + N * M + (N if vb else 0)
    S = np.zeros((n_par, n_par))
    ct = lambda w: c[w] if w >= 0 else 0.0
    for uh in range(-1, N):
        for ul in range(-1, uh + 1):
            r = _spin_row(uh, ul, N)
            rh, rl = _spin_row(uh, -1, N), _spin_row(ul, -1, N)
            for w in range(-1, M):
                pp = _param(uh, w, N, M, hb, vb)
                if pp < 0:
                    continue
                up = C[rh, _hidden_row(w, -1, M)] + (ct(w) * C[rh, 0] if w >= 0 else 0.0)
                for xx in range(-1, M):
                    qq = _param(ul, xx, N, M, hb, vb)
                    if qq < 0 or (uh == ul and pp < qq):
                        continue
                    u = C[r, _hidden_row(w, xx, M)]
                    if w >= 0:
                        u += ct(w) * C[r, _hidden_row(-1, xx, M)]
                    if xx >= 0:
                        u += ct(xx) * C[r, _hidden_row(w, -1, M)]
                    if w >= 0 and xx >= 0:
                        u += ct(w) * ct(xx) * C[r, 0]
                    uq = C[rl, _hidden_row(xx, -1, M)] + (ct(xx) * C[rl, 0] if xx >= 0 else 0.0)
                    v = scale * (u - up * uq / ns)
                    S[max(pp, qq), min(pp, qq)] = v
    return S


@pytest.mark.parametrize("N,alpha,ns,biases,bias_shift", [(5, 2, 300, (True, True), 0.0),
                                                          (6, 1, 97, (False, True), 1.5),
                                                          (4, 3, 64, (True, False), -2.0),
                                                          (5, 1, 41, (False, False), 0.7)])
def test_structured_s_equals_jacobian_product(N, alpha, ns, biases, bias_shift):
    rng = np.random.default_rng(N * 100 + ns)
    M = alpha * N
    n_par = (M if biases[0] else 0) + N * M + (N if biases[1] else 0)
    flat = rng.normal(scale=0.4, size=n_par)
    if biases[0]:
        flat[:M] += bias_shift  # activations with a large mean: the centring matters
    p = orbm.RBMParams(flat, N, M, hidden_bias=biases[0], visible_bias=biases[1])
    x = rng.choice([-1.0, 1.0], size=(ns, N))
    O = orbm.jacobian(p, x)
    Oc = O - O.mean(axis=0)
    ref = Oc.T @ Oc / ns
    S = structured_s(p, x, 1.0 / ns)
    lo = np.tril_indices(n_par)
    np.testing.assert_allclose(S[lo], ref[lo], rtol=0, atol=1e-12 * np.abs(ref).max())
    # the strict upper triangle is never written
    assert np.all(S[np.triu_indices(n_par, 1)] == 0.0)


def test_structured_s_catches_a_dropped_centring_term():
    # the same restatement with the c_w C[r,(-1,w')] term left out must fail
    rng = np.random.default_rng(5)
    N, M, ns = 4, 4, 200
    flat = rng.normal(scale=0.4, size=M + N * M + N)
    flat[:M] += 1.0
    p = orbm.RBMParams(flat, N, M)
    x = rng.choice([-1.0, 1.0], size=(ns, N))
    O = orbm.jacobian(p, x)
    Oc = O - O.mean(axis=0)
    ref = Oc.T @ Oc / ns
    good = structured_s(p, x, 1.0 / ns)
    lo = np.tril_indices(ref.shape[0])
    assert np.abs(good[lo] - ref[lo]).max() < 1e-12
    global _hidden_row
    orig = _hidden_row
    try:
        # point the (-1, w') entries at the ones row: the c_w term then adds c_w c_w' ns
        _hidden_row = lambda w, x_, M_: 0 if (w < 0 and x_ >= 0) else orig(w, x_, M_)
        bad = structured_s(p, x, 1.0 / ns)
    finally:
        _hidden_row = orig
    assert np.abs(bad[lo] - ref[lo]).max() > 1e-3
