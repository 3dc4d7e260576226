"""GPU parity: connected elements (bit-exact) and local energies vs the oracle."""

import numpy as np
import pytest
import torch

import netket_amd as nk
from oracle import operators as O
from oracle import rbm as orbm

pytestmark = pytest.mark.gpu


def _states(N, n, seed, n_down=None):
    rng = np.random.default_rng(seed)
    if n_down is None:
        return rng.choice(np.array([-1, 1], dtype=np.int8), size=(n, N))
    x = np.ones((n, N), dtype=np.int8)
    for b in range(n):
        x[b, rng.choice(N, n_down, replace=False)] = -1
    return x


def _terms(op):
    return dict(zip([tuple(a) for a in op.acting_on], op.operators))


@pytest.mark.parametrize("N,h,J,dtype", [(16, 1.0, 1.0, np.float64), (10, 0.37, -1.3, np.float64),
                                         (12, 0.5, 0.8, np.float32), (9, 0.0, 1.0, np.float64)])
def test_ising_conn_bit_exact(gpu, N, h, J, dtype):
    print(f"case ising N={N} h={h} J={J} {np.dtype(dtype)}", flush=True)
    g = nk.graph.Chain(N)
    hi = nk.hilbert.Spin(0.5, N=N)
    op = nk.operator.Ising(hi, g, h=h, J=J, dtype=dtype)
    x = _states(N, 300, 1)
    xp, mels = op.get_conn_padded(x)
    xp_o, mels_o = O.ising_conn(x, op.edges, h, J, dtype)
    np.testing.assert_array_equal(xp.cpu().numpy(), xp_o)
    assert mels.dtype == (torch.float64 if dtype == np.float64 else torch.float32)
    np.testing.assert_array_equal(mels.cpu().numpy(), mels_o)


def test_ising_conn_batch_shapes(gpu):
    # test/operator/test_operator.py:324-359: batch dims are preserved
    N = 6
    op = nk.operator.Ising(nk.hilbert.Spin(0.5, N=N), nk.graph.Chain(N), h=1.0)
    x = _states(N, 12, 2).reshape(3, 4, N)
    xp, mels = op.get_conn_padded(x)
    assert tuple(xp.shape) == (3, 4, N + 1, N) and tuple(mels.shape) == (3, 4, N + 1)
    xp2, _ = op.get_conn_padded(x.reshape(12, N))
    np.testing.assert_array_equal(xp.cpu().numpy().reshape(12, N + 1, N), xp2.cpu().numpy())


@pytest.mark.parametrize("case", ["heis_chain16", "heis_chain64_sz0", "j1j2_4x4", "mixed_terms",
                                  "heis_no_sign"])
def test_localop_conn_bit_exact(gpu, case):
    print(f"case localop {case}", flush=True)
    if case == "heis_chain16":
        N = 16
        op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N), nk.graph.Chain(N))
        x = _states(N, 400, 3)
    elif case == "heis_chain64_sz0":
        N = 64
        op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N, total_sz=0), nk.graph.Chain(N))
        x = _states(N, 257, 4, n_down=32)
    elif case == "j1j2_4x4":
        N = 16
        g = nk.graph.Square(4, max_neighbor_order=2)
        op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N, total_sz=0), g, J=[1.0, 0.5])
        x = _states(N, 300, 5, n_down=8)
    elif case == "heis_no_sign":
        N = 12
        op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N), nk.graph.Chain(N), sign_rule=False)
        x = _states(N, 300, 7)
    else:
        N = 8
        hi = nk.hilbert.Spin(0.5, N=N)
        rng = np.random.default_rng(6)
        op = nk.operator.LocalOperator(hi, constant=0.5)
        for _ in range(6):
            k = int(rng.integers(1, 4))
            s = [int(v) for v in rng.choice(N, size=k, replace=False)]
            m = rng.normal(size=(2**k, 2**k))
            m[np.abs(m) < 0.7] = 0
            op += nk.operator.LocalOperator(hi, [m + m.T], [s])
        x = _states(N, 256, 8)
    xp, mels, n = nk._ops.localop_conn(op, x, return_n_conn=True)
    xp_o, mels_o, n_o = O.local_conn(x, _terms(op), constant=float(op.constant))
    assert xp.shape[1] == xp_o.shape[1] == op.max_conn_size
    np.testing.assert_array_equal(xp.cpu().numpy(), xp_o)
    np.testing.assert_array_equal(mels.cpu().numpy(), mels_o)
    np.testing.assert_array_equal(n.cpu().numpy(), n_o)


def _rbm_state(N, alpha, dtype, seed, scale=0.1):
    model = nk.models.RBM(alpha=alpha, param_dtype=dtype)
    lay = model.layout(N)
    rng = np.random.default_rng(seed)
    flat_np = rng.normal(scale=scale, size=lay["n_params"]).astype(dtype)
    flat = torch.as_tensor(flat_np, device="cuda")
    p = orbm.RBMParams(flat_np.astype(np.float64), N, lay["M"])
    return model, flat, p


@pytest.mark.parametrize("N,alpha,dtype", [(16, 1, np.float64), (64, 2, np.float32), (20, 3, np.float64),
                                           (100, 4, np.float32)])
def test_log_psi(gpu, N, alpha, dtype):
    print(f"case log_psi N={N} alpha={alpha} {np.dtype(dtype)}", flush=True)
    model, flat, p = _rbm_state(N, alpha, dtype, 10)
    x = _states(N, 500, 11)
    lp = model.apply(flat, x).cpu().numpy()
    ref = orbm.log_psi(p, x)
    if dtype == np.float64:
        np.testing.assert_allclose(lp, ref, rtol=1e-12, atol=1e-12)
    else:
        np.testing.assert_allclose(lp, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("kind,N,alpha,dtype", [
    ("ising", 16, 1, np.float64), ("ising", 40, 2, np.float32),
    ("heis", 22, 1, np.float64), ("heis", 64, 2, np.float32),
    ("j1j2", 16, 2, np.float64), ("mixed", 8, 2, np.float64),
])
def test_local_energy(gpu, kind, N, alpha, dtype):
    print(f"case eloc {kind} N={N} alpha={alpha} {np.dtype(dtype)}", flush=True)
    model, flat, p = _rbm_state(N, alpha, dtype, 20, scale=0.2)
    if kind == "ising":
        op = nk.operator.Ising(nk.hilbert.Spin(0.5, N=N), nk.graph.Chain(N), h=0.9, J=1.1)
        x = _states(N, 300, 21)
        conn = lambda xx: O.ising_conn(xx, op.edges, 0.9, 1.1)
    else:
        if kind == "heis":
            op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N, total_sz=0), nk.graph.Chain(N))
        elif kind == "j1j2":
            op = nk.operator.Heisenberg(nk.hilbert.Spin(0.5, N=N, total_sz=0),
                                        nk.graph.Square(4, max_neighbor_order=2), J=[1.0, 0.5])
        else:
            hi = nk.hilbert.Spin(0.5, N=N)
            op = nk.operator.Ising(hi, nk.graph.Chain(N), h=0.6).to_local_operator()
            op += nk.operator.Heisenberg(hi, nk.graph.Chain(N))
        x = _states(N, 300, 22, n_down=N // 2)
        terms = _terms(op)
        conn = lambda xx: O.local_conn(xx, terms, constant=float(op.constant))
    from netket_amd import _engine as eng

    e = eng.local_energies(model, flat, op, torch.as_tensor(x, device="cuda")).cpu().numpy()
    ref = orbm.local_energy(p, x, conn)
    if dtype == np.float64:
        np.testing.assert_allclose(e, ref, rtol=1e-10, atol=1e-10)
    else:
        np.testing.assert_allclose(e, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


@pytest.mark.parametrize("lattice,alpha,scale,bias,n", [
    ("chain100", 4, 0.1, True, 300),       # M = 400: a partly empty last unit tile
    ("square10", 3, 0.3, True, 200),       # M = 300, large |theta|, 2D edges
    ("chain90", 3, 0.2, False, 131),       # no biases, N not a multiple of 8, ragged sample tile
    ("square12", 2, 0.1, True, 70),        # 144 sites: five spin words per sample
])
def test_local_energy_ising_tiled(gpu, lattice, alpha, scale, bias, n):
    # FP32 models whose W exceeds 64 KiB take the path tiled over samples and
    # hidden units (theta GEMM + flip kernel over all sites + tile reduction)
    print(f"case eloc ising tiled {lattice} alpha={alpha} bias={bias} n={n}", flush=True)
    import ctypes as C
    from netket_amd import _engine as eng, _native as nat, _ops
    g = nk.graph.Square(int(lattice[6:])) if lattice.startswith("square") else nk.graph.Chain(int(lattice[5:]))
    N = g.n_nodes
    op = nk.operator.Ising(nk.hilbert.Spin(0.5, N=N), g, h=0.8, J=0.9, dtype=np.float32)
    model = nk.models.RBM(alpha=alpha, param_dtype=np.float32, use_hidden_bias=bias,
                          use_visible_bias=bias)
    lay = model.layout(N)
    flat_np = np.random.default_rng(30 + N).normal(scale=scale, size=lay["n_params"]).astype(np.float32)
    flat = torch.as_tensor(flat_np, device="cuda")
    p = orbm.RBMParams(flat_np.astype(np.float64), N, lay["M"], hidden_bias=bias, visible_bias=bias)
    m = model.native(flat, N)
    kind, op_ptr, keep = _ops.op_struct(op)
    assert nat.lib().nka_rbm_local_energy_workspace(C.byref(m), kind, op_ptr, n) > 0
    x = _states(N, n, 31)
    ref = orbm.local_energy(p, x, lambda xx: O.ising_conn(xx, op.edges, 0.8, 0.9, np.float32))
    e, lp = eng.local_energies(model, flat, op, torch.as_tensor(x, device="cuda"), log_psi=True)
    e = e.cpu().numpy()
    np.testing.assert_allclose(e, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    np.testing.assert_allclose(lp.cpu().numpy(), orbm.log_psi(p, x), rtol=1e-5, atol=1e-5)
    # a workspace for 64 samples processes the batch in chunks, same values
    per64 = nat.lib().nka_rbm_local_energy_workspace(C.byref(m), kind, op_ptr, 64)
    ws = torch.empty(int(per64), dtype=torch.uint8, device="cuda")
    xt = torch.as_tensor(x, device="cuda")
    e2 = torch.empty(n, dtype=torch.float64, device="cuda")
    nat.call("nka_rbm_local_energy", C.byref(m), kind, op_ptr, _ops.ptr(xt), n, _ops.ptr(e2), None,
             _ops.ptr(ws), ws.numel(), _ops.stream())
    np.testing.assert_array_equal(e2.cpu().numpy(), e)
    # without a workspace the call is refused
    with pytest.raises(nat.NativeError, match="workspace"):
        nat.call("nka_rbm_local_energy", C.byref(m), kind, op_ptr, _ops.ptr(xt), n, _ops.ptr(e2), None,
                 None, 0, _ops.stream())
    del keep
