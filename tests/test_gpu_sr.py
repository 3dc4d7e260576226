"""GPU parity of statistics, energy gradient, QGT/NTK Gram matrices and the
dense solvers vs the oracle, on identical sample sets."""

import math

import numpy as np
import pytest
import scipy.linalg as sla
import torch

import netket_amd as nk
from netket_amd import _engine as eng
from netket_amd import _native as nat
from netket_amd import _ops
from netket_amd.optimizer import solver as gsolver
from oracle import ngd as ongd
from oracle import operators as O
from oracle import rbm as orbm
from oracle import stats as ostats

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C,L", [(16, 63), (4096, 4), (4, 256), (4, 4), (1, 100), (33, 7)])
def test_statistics(gpu, C, L):
    print(f"case stats {C}x{L}", flush=True)
    rng = np.random.default_rng(C * 1000 + L)
    data = rng.normal(loc=-3.0, size=(C, L)) + 0.3 * np.cumsum(rng.normal(size=(C, L)), axis=1)
    st = nk.stats.statistics(torch.as_tensor(data, device="cuda"))
    ref = ostats.statistics(data)
    for k, v in (("mean", st.mean), ("variance", st.variance), ("error_of_mean", st.error_of_mean),
                 ("tau_corr", st.tau_corr), ("R_hat", st.R_hat)):
        if np.isnan(ref[k]):
            assert np.isnan(v), k
        else:
            assert v == pytest.approx(ref[k], rel=1e-10, abs=1e-12), k


def _state(kind, dtype, n_chains, n_samples, seed=0, scale=0.05):
    if kind == "ising":
        N = 12
        hi = nk.hilbert.Spin(0.5, N=N)
        H = nk.operator.Ising(hi, nk.graph.Chain(N), h=1.0)
        sa = nk.sampler.MetropolisLocal(hi, n_chains=n_chains)
        conn = lambda x: O.ising_conn(x, H.edges, 1.0, 1.0)
    else:
        N = 16
        hi = nk.hilbert.Spin(0.5, N=N, total_sz=0)
        H = nk.operator.Heisenberg(hi, nk.graph.Chain(N))
        sa = nk.sampler.MetropolisExchange(hi, graph=nk.graph.Chain(N), n_chains=n_chains)
        terms = dict(zip([tuple(a) for a in H.acting_on], H.operators))
        conn = lambda x: O.local_conn(x, terms)
    model = nk.models.RBM(alpha=2, param_dtype=dtype)
    vs = nk.vqs.MCState(sa, model, n_samples=n_samples, n_discard_per_chain=4, seed=seed)
    # larger random parameters than the default init so the gradient is not tiny
    rng = np.random.default_rng(seed + 1)
    flat = rng.normal(scale=scale, size=vs.n_parameters).astype(dtype)
    vs.parameters = torch.as_tensor(flat, device="cuda")
    p = orbm.RBMParams(flat.astype(np.float64), N, model.n_hidden(N))
    return vs, H, p, conn


@pytest.mark.parametrize("kind,dtype", [("ising", np.float64), ("ising", np.float32),
                                        ("heis", np.float64), ("heis", np.float32)])
def test_expect_and_grad(gpu, kind, dtype):
    print(f"case expect_and_grad {kind} {np.dtype(dtype)}", flush=True)
    vs, H, p, conn = _state(kind, dtype, 32, 2048, scale=0.1)
    st, grad = vs.expect_and_grad(H)
    samples = vs.samples.cpu().numpy()
    ref_st, ref_g, _ = orbm.expect_and_grad(p, samples, conn)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    assert st.mean == pytest.approx(ref_st["mean"], rel=tol)
    assert st.variance == pytest.approx(ref_st["variance"], rel=10 * tol)
    g = vs.flatten_tree(grad).cpu().numpy()
    np.testing.assert_allclose(g, ref_g, rtol=tol, atol=tol * np.abs(ref_g).max())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_qgt_dense(gpu, dtype):
    print(f"case qgt {np.dtype(dtype)}", flush=True)
    vs, H, p, conn = _state("ising", dtype, 16, 1008, scale=0.2)
    qgt = nk.optimizer.QGTJacobianDense(vs, diag_shift=0.01)
    S = qgt.to_dense().cpu().numpy()
    x = vs.samples.cpu().numpy().reshape(-1, vs.hilbert.size)
    OL = ongd.centred_jacobian(orbm.jacobian(p, x))
    S_ref = ongd.qgt_dense(OL, 0.01)
    np.testing.assert_array_equal(S, S.T)
    if dtype == np.float64:
        np.testing.assert_allclose(S, S_ref, rtol=1e-12, atol=1e-14)
    else:
        # FP32 RBM: S from the spin-pair x centred-activation-pair product
        # (bf16 hi + lo keeps ~16 mantissa bits of each activation pair, the
        # accumulation is chunked); the same bound as for the bf16x3 Gram of
        # the Jacobian, |dS_pq| <= c 2^-16 sqrt(S_pp S_qq)
        d = np.sqrt(np.diag(S_ref - 0.01 * np.eye(len(S_ref))))
        bound = 3e-5 * np.outer(d, d) + 1e-9
        err = np.abs(S - S_ref)
        print(f"max |dS|/bound = {(err / bound).max():.3f}", flush=True)
        assert np.all(err <= bound)


@pytest.mark.parametrize("dtype,ns,biases,N", [(np.float64, 300, (True, True), 14),
                                               (np.float64, 64, (False, True), 14),
                                               (np.float64, 129, (True, False), 14),
                                               (np.float32, 300, (True, True), 14),
                                               (np.float32, 129, (False, True), 70),
                                               (np.float32, 260, (True, False), 97),
                                               (np.float32, 200, (True, True), 150),
                                               (np.float64, 1, (True, True), 14)])
def test_rbm_ntk_structured(gpu, dtype, ns, biases, N):
    # T = (O - mean)(O - mean)^T / ns + shift I + offset 11^T from the RBM
    # structure equals the explicit product of the oracle's FP64 Jacobian
    # (netket/_src/ngd/srt.py:55-66 with proj_reg giving the offset)
    print(f"case rbm_ntk {np.dtype(dtype)} ns={ns} biases={biases} N={N}", flush=True)
    alpha = 3
    M = alpha * N
    model = nk.models.RBM(alpha=alpha, param_dtype=dtype, use_hidden_bias=biases[0],
                          use_visible_bias=biases[1])
    rng = np.random.default_rng(ns + 7)
    n_par = model.layout(N)["n_params"]
    flat = rng.normal(scale=0.3, size=n_par).astype(dtype)
    x = rng.choice(np.array([-1, 1], dtype=np.int8), size=(ns, N))
    sig = torch.as_tensor(x, device="cuda")
    fl = torch.as_tensor(flat, device="cuda")
    t = eng.prepare(model, fl, sig)[0]
    ones = torch.full((1, ns), 1.0 / ns, dtype=torch.float64, device="cuda")
    mean = eng.contract(model, fl, sig, t, ones)[0]
    shift, offset = 1e-3, 0.25 / ns
    T = eng.rbm_ntk(model, fl, sig, t, mean, 1.0 / ns, shift, offset).cpu().numpy()
    p = orbm.RBMParams(flat.astype(np.float64), N, M, hidden_bias=biases[0],
                       visible_bias=biases[1])
    Ou = orbm.jacobian(p, x)
    Oc = Ou - Ou.mean(axis=0)
    ref = Oc @ Oc.T / ns + shift * np.eye(ns) + offset
    lo = np.tril_indices(ns)
    # the centring is applied to the uncentred products, so rounding scales
    # with the uncentred norms |O_s| |O_t| / ns
    dc = np.sqrt(np.sum(Oc * Oc, axis=1) / ns)
    du = np.sqrt(np.sum(Ou * Ou, axis=1) / ns)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    # FP32 models: t.t on bf16x3 keeps ~16 mantissa bits of t per product, so
    # |dTheta_st| <= c 2^-16 |t_s||t_t| and |dT_st| <= scale (hb + N) |dTheta_st|
    # <= c 2^-16 du_s du_t; sigma.sigma and the centring are exact / FP64
    tol_u = 1e-14 if dtype == np.float64 else 3e-5
    bound = (tol * np.outer(dc, dc) + tol_u * np.outer(du, du))[lo] + 1e-15
    err = np.abs(T[lo] - ref[lo])
    print(f"max |dT| / bound = {(err / bound).max():.2e}", flush=True)
    assert np.all(err <= bound)


@pytest.mark.parametrize("ns,biases,N,alpha", [(300, (True, True), 14, 3),
                                               (129, (False, True), 14, 3),
                                               (64, (True, False), 16, 2),
                                               (1, (True, True), 8, 1),
                                               (1000, (False, False), 10, 2),
                                               (257, (True, True), 70, 3),
                                               (4099, (True, True), 32, 2)])
def test_rbm_qgt_structured(gpu, ns, biases, N, alpha):
    # S = (O - mean)^T (O - mean) / ns + shift I from the RBM's spin-pair x
    # activation-pair products equals the explicit product of the oracle's
    # FP64 Jacobian (netket/_src/ngd/sr.py:50-55).  M = 210 at N = 70 puts a
    # C row (22366 floats) past the LDS staging of the unpack.
    print(f"case rbm_qgt ns={ns} biases={biases} N={N} alpha={alpha}", flush=True)
    M = alpha * N
    model = nk.models.RBM(alpha=alpha, param_dtype=np.float32, use_hidden_bias=biases[0],
                          use_visible_bias=biases[1])
    rng = np.random.default_rng(ns + N)
    n_par = model.layout(N)["n_params"]
    flat = rng.normal(scale=0.3, size=n_par).astype(np.float32)
    x = rng.choice(np.array([-1, 1], dtype=np.int8), size=(ns, N))
    sig = torch.as_tensor(x, device="cuda")
    fl = torch.as_tensor(flat, device="cuda")
    t, tT = eng.prepare(model, fl, sig, row_major=True, transposed=True)
    shift = 1e-2
    out = torch.full((n_par, n_par), 7.0, dtype=torch.float64, device="cuda")
    S = eng.qgt_structured(model, fl, sig, tT, 1.0 / ns, shift, out=out).cpu().numpy()
    p = orbm.RBMParams(flat.astype(np.float64), N, M, hidden_bias=biases[0], visible_bias=biases[1])
    Ou = orbm.jacobian(p, x)
    Oc = Ou - Ou.mean(axis=0)
    ref = Oc.T @ Oc / ns + shift * np.eye(n_par)
    lo = np.tril_indices(n_par)
    # the centring is applied to the uncentred products: rounding scales with
    # the uncentred column norms.  Each activation pair keeps ~16 mantissa bits
    # (bf16 hi + lo), the spin pairs are exact, accumulation is FP32.
    du = np.sqrt(np.sum(Ou * Ou, axis=0) / ns)
    bound = (3e-5 * np.outer(du, du))[lo] + 1e-12
    err = np.abs(S[lo] - ref[lo])
    print(f"max |dS| / bound = {(err / bound).max():.2e}", flush=True)
    assert np.all(err <= bound)
    # the strict upper triangle is not written
    up = np.triu_indices(n_par, 1)
    assert np.all(S[up] == 7.0)


def test_rbm_qgt_structured_large_mean_activations(gpu):
    # hidden biases of 2 put every tanh near 0.96 with a small spread, so
    # E[t_j t_l] is far larger than S's entries: the activations are centred
    # before the bf16 split, and the error must stay small against the
    # centred diagonal, not only against the uncentred norms
    print("case rbm_qgt large-mean activations", flush=True)
    N, alpha, ns = 16, 2, 2048
    M = alpha * N
    model = nk.models.RBM(alpha=alpha, param_dtype=np.float32)
    lay = model.layout(N)
    rng = np.random.default_rng(11)
    flat = rng.normal(scale=0.05, size=lay["n_params"]).astype(np.float32)
    b0 = lay["bias"][0]
    flat[b0:b0 + M] = 2.0
    x = rng.choice(np.array([-1, 1], dtype=np.int8), size=(ns, N))
    sig = torch.as_tensor(x, device="cuda")
    fl = torch.as_tensor(flat, device="cuda")
    t, tT = eng.prepare(model, fl, sig, row_major=True, transposed=True)
    S = eng.qgt_structured(model, fl, sig, tT, 1.0 / ns, 0.0).cpu().numpy()
    p = orbm.RBMParams(flat.astype(np.float64), N, M)
    Ou = orbm.jacobian(p, x)
    Oc = Ou - Ou.mean(axis=0)
    ref = Oc.T @ Oc / ns
    lo = np.tril_indices(lay["n_params"])
    d = np.sqrt(np.diag(ref))
    ratio = (np.abs(S - ref) / np.outer(d, d))[lo].max()
    spread = (np.sqrt(np.sum(Ou * Ou, axis=0) / ns) / d)[b0:b0 + M].min()
    print(f"uncentred / centred norm of the hidden-bias columns >= {spread:.1f}; "
          f"max |dS| / sqrt(S_ii S_jj) = {ratio:.2e}", flush=True)
    assert spread > 20.0
    assert ratio < 3e-5


def test_rbm_qgt_structured_full_size(gpu):
    # the benchmark's shape (N = 64, alpha = 2, 16384 samples, S 8384^2) against
    # the FP64 product of the same Jacobian formed on the GPU
    print("case rbm_qgt full size", flush=True)
    N, alpha, ns = 64, 2, 16384
    model = nk.models.RBM(alpha=alpha, param_dtype=np.float32)
    rng = np.random.default_rng(3)
    n_par = model.layout(N)["n_params"]
    flat = rng.normal(scale=0.05, size=n_par).astype(np.float32)
    x = np.ones((ns, N), dtype=np.int8)
    for s in range(ns):
        x[s, rng.permutation(N)[: N // 2]] = -1
    sig = torch.as_tensor(x, device="cuda")
    fl = torch.as_tensor(flat, device="cuda")
    t, tT = eng.prepare(model, fl, sig, row_major=True, transposed=True)
    S = eng.qgt_structured(model, fl, sig, tT, 1.0 / ns, 0.01)
    lay = model.layout(N)
    M = lay["M"]
    t64 = t.double()
    s64 = sig.double()
    O = torch.empty((ns, n_par), dtype=torch.float64, device="cuda")
    O[:, lay["bias"][0]: lay["bias"][0] + M] = t64
    O[:, lay["kernel"][0]: lay["kernel"][0] + N * M] = (s64[:, :, None] * t64[:, None, :]).reshape(ns, N * M)
    O[:, lay["visible_bias"][0]: lay["visible_bias"][0] + N] = s64
    du = torch.sqrt((O * O).sum(0) / ns)
    O -= O.mean(0)
    ref = O.T @ O / ns
    del O
    d = torch.sqrt(ref.diagonal().clone())
    ref.diagonal().add_(0.01)
    L = torch.tril(torch.ones(n_par, n_par, dtype=torch.bool, device="cuda"))
    err = (S - ref).abs()
    bound = 3e-5 * du[:, None] * du[None, :] + 1e-12
    ratio = torch.where(L, err / bound, torch.zeros_like(err)).max().item()
    # the centred activations, the column sums taken from the same product
    # and the chunked accumulation keep the error small against the centred
    # diagonal too
    rel = torch.where(L, err / (d[:, None] * d[None, :]), torch.zeros_like(err)).max().item()
    print(f"max |dS| / bound = {ratio:.2e}; max |dS| / sqrt(S_ii S_jj) = {rel:.2e}", flush=True)
    assert ratio <= 1.0
    assert rel < 3e-5


@pytest.mark.parametrize("solver_name", ["cholesky", "cg"])
def test_qgt_diag_scale(gpu, solver_name):
    # scale-invariant regularisation (netket/optimizer/qgt/qgt_jacobian_common.py:21-60,
    # qgt_jacobian_dense.py:93-161): the solution satisfies
    # (S + diag_scale diag(S) + diag_shift I) x = F, and to_dense is that matrix
    print(f"case qgt diag_scale {solver_name}", flush=True)
    vs, H, p, conn = _state("ising", np.float64, 16, 1008, scale=0.2)
    shift, dscale = 0.01, 0.1
    qgt = nk.optimizer.QGTJacobianDense(vs, diag_shift=shift, diag_scale=dscale)
    x = vs.samples.cpu().numpy().reshape(-1, vs.hilbert.size)
    OL = ongd.centred_jacobian(orbm.jacobian(p, x))
    S = OL.T @ OL
    A = S + dscale * np.diag(np.diag(S)) + shift * np.eye(len(S))
    np.testing.assert_allclose(qgt.to_dense().cpu().numpy(), A, rtol=1e-12, atol=1e-14)
    F = np.random.default_rng(7).normal(size=len(S))
    fun = gsolver.cholesky if solver_name == "cholesky" else gsolver.cg
    kw = {} if solver_name == "cholesky" else {"tol": 1e-12}
    sol, _ = qgt.solve(lambda A_, b_, x0=None: fun(A_, b_, x0=x0, **kw), torch.as_tensor(F, device="cuda"))
    ref = np.linalg.solve(A, F)
    np.testing.assert_allclose(sol.cpu().numpy(), ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    if solver_name == "cg":
        # the CG sees the rescaled system: the iterations of the oracle's CG on
        # D^-1 S D^-1 + scale I, not those on S + scale diag(S) + shift I (which
        # needs many more); one either way for the last-bit differences of S
        d = np.sqrt(np.diag(S) + shift / dscale)
        Ss = S / np.outer(d, d) + dscale * np.eye(len(S))
        ys, k = ongd.cg(lambda v: Ss @ v, F / d, tol=1e-6)
        _, k_unscaled = ongd.cg(lambda v: A @ v, F, tol=1e-6)
        _, info = qgt.solve(lambda A_, b_, x0=None: gsolver.cg(A_, b_, x0=x0, tol=1e-6),
                            torch.as_tensor(F, device="cuda"))
        print(f"CG iterations: {info['iterations']} (oracle rescaled {k}, unscaled {k_unscaled})", flush=True)
        assert abs(info["iterations"] - k) <= 1


def test_cg_from_x0(gpu):
    # jax.scipy.sparse.linalg.cg from a given x0: r0 = b - A x0, tolerance relative to b.b;
    # SR with solver_restart=True passes the previous solution
    n = 150
    rng = np.random.default_rng(12)
    A = rng.normal(size=(n, n))
    A = A @ A.T / n + 0.5 * np.eye(n)
    b = rng.normal(size=n)
    x0 = np.linalg.solve(A, b) + 1e-3 * rng.normal(size=n)
    x, info = gsolver.cg(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"),
                         x0=torch.as_tensor(x0, device="cuda"))
    ref, k = ongd.cg(lambda v: A @ v, b, x0=x0)
    x_cold, k_cold = ongd.cg(lambda v: A @ v, b)
    assert info["iterations"] == k and k < k_cold
    np.testing.assert_allclose(x.cpu().numpy(), ref, rtol=1e-8, atol=1e-10)
    # an x0 that already solves the system: no iteration
    xs = np.linalg.solve(A, b)
    _, info0 = gsolver.cg(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"),
                          x0=torch.as_tensor(xs, device="cuda"))
    assert info0["iterations"] == 0
    # through SR with solver_restart=True: the second solve starts from the first solution
    vs, H, p, conn = _state("ising", np.float64, 16, 1008, scale=0.2)
    sr = nk.optimizer.SR(nk.optimizer.QGTJacobianDense, solver=gsolver.cg, diag_shift=0.01,
                         solver_restart=True)
    st, g = vs.expect_and_grad(H)
    x1 = sr(vs, g)
    it1 = sr.info["iterations"]
    sr(vs, g)
    assert sr.info["iterations"] < it1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_prepare_large_model(gpu, dtype):
    # W larger than the LDS budget takes the tiled matrix-product path
    print(f"case prepare large {np.dtype(dtype)}", flush=True)
    N, alpha, ns = 70, 5, 333
    model = nk.models.RBM(alpha=alpha, param_dtype=dtype)
    M = model.n_hidden(N)
    flat = np.random.default_rng(3).normal(scale=0.2, size=model.layout(N)["n_params"]).astype(dtype)
    x = np.random.default_rng(4).choice(np.array([-1, 1], dtype=np.int8), size=(ns, N))
    t, tT = eng.prepare(model, torch.as_tensor(flat, device="cuda"), torch.as_tensor(x, device="cuda"),
                        row_major=True, transposed=True)
    p = orbm.RBMParams(flat.astype(np.float64), N, M)
    ref = np.tanh(orbm.theta(p, x))
    tol = 2e-6 if dtype == np.float32 else 1e-13
    np.testing.assert_allclose(t.cpu().numpy(), ref, rtol=0, atol=tol)
    np.testing.assert_allclose(tT.cpu().numpy(), ref.T, rtol=0, atol=tol)


@pytest.mark.parametrize("n", [1, 63, 64, 200, 333, 777, 1500])
def test_cholesky_solve(gpu, n):
    print(f"case cholesky n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    A = A @ A.T / n + 0.05 * np.eye(n)
    b = rng.normal(size=n)
    x, info = gsolver.cholesky(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"))
    ref = sla.cho_solve(sla.cho_factor(A), b)
    np.testing.assert_allclose(x.cpu().numpy(), ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("n,lda", [(769, 769), (1536, 1536), (1600, 1601), (2047, 2048), (2560, 2561),
                                   (4192, 4192), (8384, 8384), (16382, 16382)])
def test_cholesky_many_panels(gpu, n, lda):
    # several panels (768 columns below n = 16384), a narrow last panel (one
    # column at n = 769), even and odd row strides and padded rows: L L^T
    # against A on sampled rows, the solve's residual, and L against
    # rocSOLVER's factor
    print(f"case cholesky panels n={n} lda={lda}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(n)
    X = torch.randn(n, n // 2 + 7, dtype=torch.float64, device="cuda", generator=g)
    A0 = X @ X.T / n + 0.05 * torch.eye(n, dtype=torch.float64, device="cuda")
    del X
    buf = torch.full((n, lda), float("nan"), dtype=torch.float64, device="cuda")
    buf[:, :n] = A0
    a = buf[:, :n]
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = _ops.potrf_(a, info)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    _ops.potrs_(a, ws, b, info)
    assert int(info.item()) == 0
    assert (A0 @ b - 1.0).abs().max().item() < 1e-9
    L = torch.tril(a)
    Lref = torch.linalg.cholesky(A0)
    assert (L - Lref).abs().max().item() < 2e-9 * Lref.abs().max().item()
    rows = torch.arange(0, n, 97, device="cuda")
    LLt = L[rows] @ L[rows].T
    assert (LLt - A0[rows][:, rows]).abs().max().item() < 1e-11
    if lda > n:   # the padding columns are never written
        assert torch.isnan(buf[:, n:]).all()
    del a, A0, ws, buf, L, Lref


def test_cholesky_large_many_panel_workgroups(gpu):
    # n > 512 * 64: the first panel launches more workgroups than the chip
    # holds at once, so some start after workgroup 0 has factored the diagonal
    # block; the factor and the solve must not depend on that order.  Checked
    # by the residual of the solve and by L L^T on a sample of rows/columns.
    n = 33000
    print(f"case cholesky large n={n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    X = torch.randn(n, n // 2, dtype=torch.float64, device="cuda", generator=g)
    A0 = X @ X.T / n + 0.1 * torch.eye(n, dtype=torch.float64, device="cuda")
    del X
    a = A0.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = _ops.potrf_(a, info)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    _ops.potrs_(a, ws, b, info)
    assert int(info.item()) == 0
    assert (A0 @ b - 1.0).abs().max().item() < 1e-9
    rows = torch.arange(0, n, 331, device="cuda")
    cols = torch.arange(n, device="cuda")
    Ls = torch.where(cols[None, :] <= rows[:, None], a[rows], 0.0)   # sampled rows of L
    LLt = Ls @ Ls.T
    ref = A0[rows][:, rows]
    assert (LLt - ref).abs().max().item() < 1e-11
    del a, A0, ws


def test_cholesky_failure_and_fallback(gpu):
    # netket/_src/solvers/nan_fallback.py: NaN output triggers pinv_smooth.
    # The test is made on the device: the primary solution comes back zeroed
    # and the flag set, with no host synchronisation inside the solver.
    n = 96
    rng = np.random.default_rng(3)
    U = rng.normal(size=(n, 40))
    A = U @ U.T  # rank 40: not positive definite
    A[70, 70] = -1.0
    b = rng.normal(size=n)
    x, _ = gsolver.cholesky(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"))
    assert np.all(np.isnan(x.cpu().numpy()))
    x2, info = gsolver.cholesky_with_fallback(torch.as_tensor(A, device="cuda"),
                                              torch.as_tensor(b, device="cuda"))
    fb = info["solver_fallback"]
    assert isinstance(fb, gsolver.DeviceFallback)
    assert int(fb.flag.item()) != 0 and bool(fb)
    assert np.all(x2.cpu().numpy() == 0.0)
    np.testing.assert_allclose(fb.solve().cpu().numpy(), ongd.pinv_smooth(A, b), rtol=1e-7, atol=1e-8)
    # a NaN right-hand side also flags
    bn = b.copy()
    bn[5] = np.nan
    B = U @ U.T + np.eye(n) * 5.0
    _, info_n = gsolver.cholesky_with_fallback(torch.as_tensor(B, device="cuda"),
                                               torch.as_tensor(bn, device="cuda"))
    assert bool(info_n["solver_fallback"])
    # a good system does not fall back and is solved by the Cholesky
    x3, info3 = gsolver.cholesky_with_fallback(torch.as_tensor(B, device="cuda"),
                                               torch.as_tensor(b, device="cuda"))
    assert not bool(info3["solver_fallback"])
    np.testing.assert_allclose(x3.cpu().numpy(), np.linalg.solve(B, b), rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("use_ntk", [False, True])
def test_vmc_sr_fallback_resolved_by_driver(gpu, use_ntk):
    # A diag_shift that makes the system indefinite: the device flag fires,
    # the driver's eager step replaces the zeroed solution by pinv_smooth's,
    # so the update equals the oracle's pinv_smooth update; a captured step
    # records the same solver with no host read and applies no update.
    print(f"case vmc_sr fallback use_ntk={use_ntk}", flush=True)
    vs, H, p, conn = _state("ising", np.float64, 8, 64, scale=0.1)
    samples = vs.samples.cpu().numpy()
    x = samples.reshape(-1, samples.shape[-1])
    eps = -1e3  # S + eps I (T + eps I) is negative definite: the factorisation fails
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=eps, variational_state=vs,
                    use_ntk=use_ntk, linear_solver=gsolver.cholesky_with_fallback)
    st, dp = drv.compute_loss_and_update()
    assert isinstance(drv.info["solver_fallback"], gsolver.DeviceFallback)
    dp = drv.resolve_fallback(dp)
    assert drv.info["solver_fallback"] is True
    e = orbm.local_energy(p, x, conn)
    OL, dv = ongd.prepare(orbm.jacobian(p, x), e)
    if use_ntk:
        T = OL @ OL.T + eps * np.eye(OL.shape[0])
        ref = OL.T @ ongd.pinv_smooth(T, dv)
    else:
        S = OL.T @ OL + eps * np.eye(OL.shape[1])
        ref = ongd.pinv_smooth(S, OL.T @ dv)
    got = dp.cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


def test_cg(gpu):
    n = 150
    rng = np.random.default_rng(11)
    A = rng.normal(size=(n, n))
    A = A @ A.T / n + 0.5 * np.eye(n)
    b = rng.normal(size=n)
    x, info = gsolver.cg(torch.as_tensor(A, device="cuda"), torch.as_tensor(b, device="cuda"))
    ref, k = ongd.cg(lambda v: A @ v, b)
    assert info["iterations"] == k
    np.testing.assert_allclose(x.cpu().numpy(), ref, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ntk_method", ["structured", "jacobian"])
def test_vmc_sr_update_matches_oracle_and_minsr(gpu, dtype, ntk_method):
    # VMC_SR with use_ntk=False and use_ntk=True give the same update
    # (test/driver/test_srt.py:98-211), and both match the oracle's SR update
    print(f"case sr/minsr {np.dtype(dtype)} {ntk_method}", flush=True)
    vs, H, p, conn = _state("ising", dtype, 16, 256, scale=0.1)
    samples = vs.samples.cpu().numpy()
    x = samples.reshape(-1, vs.hilbert.size)
    e = orbm.local_energy(p, x, conn)
    OL, dv = ongd.prepare(orbm.jacobian(p, x), e)
    ref_sr, _, _ = ongd.sr_update(OL, dv, 1e-3)
    ref_srt, _, _ = ongd.srt_update(OL, dv, 1e-3)
    np.testing.assert_allclose(ref_sr, ref_srt, rtol=1e-8, atol=1e-10)
    out = {}
    for ntk in (False, True):
        drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.01), diag_shift=1e-3, variational_state=vs,
                        use_ntk=ntk, linear_solver=gsolver.cholesky, ntk_method=ntk_method)
        st, dp = drv.compute_loss_and_update()
        out[ntk] = dp.cpu().numpy()
        assert vs.samples.cpu().numpy() is not None
    scale = np.abs(ref_sr).max()
    tol = 1e-8 if dtype == np.float64 else 2e-3
    print(f"max |sr - ref| / max|ref| = {np.abs(out[False] - ref_sr).max() / scale:.2e}, "
          f"ntk: {np.abs(out[True] - ref_sr).max() / scale:.2e}", flush=True)
    np.testing.assert_allclose(out[False], ref_sr, rtol=0, atol=tol * scale)
    np.testing.assert_allclose(out[True], ref_sr, rtol=0, atol=tol * scale)


@pytest.mark.parametrize("use_ntk,ntk_method", [(False, "structured"), (True, "structured"),
                                                (True, "jacobian")])
def test_vmc_sr_spring_momentum_two_steps(gpu, use_ntk, ntk_method):
    # SPRING (momentum 0.8): two chained steps against the oracle's SR / minSR
    # momentum formulas on the same samples (netket/_src/ngd/sr.py:43-68, srt.py:34-82)
    print(f"case spring momentum use_ntk={use_ntk} {ntk_method}", flush=True)
    vs, H, p, conn = _state("ising", np.float64, 16, 256, scale=0.1)
    N, M = vs.hilbert.size, vs.model.n_hidden(vs.hilbert.size)
    drv = nk.VMC_SR(H, nk.optimizer.Sgd(0.05), diag_shift=1e-2, variational_state=vs, use_ntk=use_ntk,
                    momentum=0.8, linear_solver=gsolver.cholesky, ntk_method=ntk_method)
    old = np.zeros(vs.n_parameters)
    for step in range(2):
        drv.reset_step()
        flat = vs.parameters_flat.cpu().numpy()
        x = vs.samples.cpu().numpy().reshape(-1, N)
        st, dp = drv.compute_loss_and_update()
        pp = orbm.RBMParams(flat, N, M)
        OL, dv = ongd.prepare(orbm.jacobian(pp, x), orbm.local_energy(pp, x, conn))
        if use_ntk:
            ref, _, _ = ongd.srt_update(OL, dv, 1e-2, momentum=0.8, old_updates=old)
        else:
            ref, _, _ = ongd.sr_update(OL, dv, 1e-2, momentum=0.8, old_updates=old)
        got = dp.cpu().numpy()
        print(f"step {step}: max|dp - ref| / max|ref| = {np.abs(got - ref).max() / np.abs(ref).max():.2e}",
              flush=True)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-8 * np.abs(ref).max())
        old = ref
        drv.update_parameters(dp)
        drv._step_count += 1


def test_vmc_sr_preconditioner_step(gpu):
    # VMC + SR(QGTJacobianDense, cholesky): dp = (S + eps I)^-1 grad on the same samples
    vs, H, p, conn = _state("heis", np.float64, 16, 512, scale=0.1)
    sr = nk.optimizer.SR(nk.optimizer.QGTJacobianDense, solver=gsolver.cholesky, diag_shift=0.01)
    drv = nk.VMC(H, nk.optimizer.Sgd(0.01), variational_state=vs, preconditioner=sr)
    st, dp = drv.compute_loss_and_update()
    x = vs.samples.cpu().numpy().reshape(-1, vs.hilbert.size)
    e = orbm.local_energy(p, x, conn)
    OL, dv = ongd.prepare(orbm.jacobian(p, x), e)
    ref, _, _ = ongd.sr_update(OL, dv, 0.01)
    np.testing.assert_allclose(dp.cpu().numpy(), ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    # the default SR solver (CG) converges to the same update within its tolerance
    sr_cg = nk.optimizer.SR(nk.optimizer.QGTJacobianDense, diag_shift=0.01)
    g = vs.expect_and_forces_flat(H)[1]  # same cached samples
    dp2 = sr_cg(vs, g)
    assert np.linalg.norm(dp2.cpu().numpy() - ref) <= 1e-3 * np.linalg.norm(ref)
