"""Pins of the FFNN oracle (CPU): the chain-rule Jacobian against central
finite differences, special cases of the forward pass, and the structured
form of the centred NTK used by the GPU path."""

import numpy as np

from oracle import ffnn as of
from oracle import rbm as orbm


def _params(N, H1, H2, seed, scale=0.4, use_bias=True):
    rng = np.random.default_rng(seed)
    n = (H1 if use_bias else 0) + N * H1 + (H2 if use_bias else 0) + H1 * H2
    flat = rng.normal(scale=scale, size=n)
    return flat, of.FFNNParams(flat, N, H1, H2, use_bias)


def test_log_cosh_against_numpy_cosh():
    x = np.linspace(-20, 20, 401)
    np.testing.assert_allclose(orbm.log_cosh(x), np.log(np.cosh(x)), rtol=1e-14, atol=1e-14)


def test_forward_special_cases():
    N, H1, H2 = 6, 5, 4
    flat, p = _params(N, H1, H2, 0)
    x = np.random.default_rng(1).choice([-1.0, 1.0], size=(9, N))
    # W2 = 0: log psi = sum_k log cosh(b2_k) for every configuration
    p.W2[:] = 0.0
    np.testing.assert_allclose(of.log_psi(p, x), np.sum(np.log(np.cosh(p.b2))), rtol=1e-14)
    # b1 = 0: log cosh is even, so log psi(x) = log psi(-x)
    flat, p = _params(N, H1, H2, 2)
    p.b1[:] = 0.0
    np.testing.assert_allclose(of.log_psi(p, x), of.log_psi(p, -x), rtol=1e-13)
    # one unit per layer: log cosh(w2 log cosh(w1.x + b1) + b2)
    flat, p = _params(N, 1, 1, 3)
    ref = np.log(np.cosh(p.W2[0, 0] * np.log(np.cosh(x @ p.W1[:, 0] + p.b1[0])) + p.b2[0]))
    np.testing.assert_allclose(of.log_psi(p, x), ref, rtol=1e-13)


def test_jacobian_matches_finite_differences():
    for use_bias in (True, False):
        N, H1, H2 = 5, 4, 3
        flat, p = _params(N, H1, H2, 4, use_bias=use_bias)
        x = np.random.default_rng(5).choice([-1.0, 1.0], size=(7, N))
        O = of.jacobian(p, x)
        O_fd = of.jacobian_fd(p, flat, x, use_bias=use_bias)
        assert O.shape == (7, p.n_params)
        np.testing.assert_allclose(O, O_fd, rtol=1e-7, atol=1e-8)


def test_structured_ntk_identity():
    # (O O^T)_st = (1 + h1_s.h1_t)(g2_s.g2_t) + (1 + x_s.x_t)(d1_s.d1_t), the form the
    # GPU kernel uses, equals the explicit product of the Jacobian
    N, H1, H2 = 6, 5, 4
    flat, p = _params(N, H1, H2, 6)
    x = np.random.default_rng(7).choice([-1.0, 1.0], size=(11, N))
    O = of.jacobian(p, x)
    th1, h1, z2 = of.forward(p, x)
    g2 = np.tanh(z2)
    d1 = (g2 @ p.W2.T) * np.tanh(th1)
    T = (1 + h1 @ h1.T) * (g2 @ g2.T) + (1 + x @ x.T) * (d1 @ d1.T)
    np.testing.assert_allclose(T, O @ O.T, rtol=1e-12, atol=1e-12)
