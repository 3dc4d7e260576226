"""The allocation-free Cholesky solve (nka_cholesky_solve_f64_noalloc) against
the workspace pair nka_potrf_f64 + nka_potrs_f64, through the C ABI, on
consecutive calls whose sizes shrink (the static scratch keeps the previous
call's contents), and on a matrix that is not positive definite."""

import ctypes as C

import numpy as np
import pytest
import torch

from netket_amd import _native as nat

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _spd(n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    X = torch.randn(n, n // 2 + 8, dtype=torch.float64, device="cuda", generator=g)
    A = X @ X.T / n + 0.05 * torch.eye(n, dtype=torch.float64, device="cuda")
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    return A, b


def _with_workspace(A, b):
    lib = nat.lib()
    n = A.shape[0]
    a = A.clone()
    x = b.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ws = torch.empty(lib.nka_potrf_workspace(n), dtype=torch.uint8, device="cuda")
    nat.check(lib.nka_potrf_f64(_ptr(a), n, n, _ptr(ws), _ptr(info), _stream()), "nka_potrf_f64")
    nat.check(lib.nka_potrs_f64(_ptr(a), n, n, _ptr(ws), _ptr(x), _ptr(info), _stream()), "nka_potrs_f64")
    return a, x, info


def _no_alloc(A, b):
    lib = nat.lib()
    n = A.shape[0]
    a = A.clone()
    x = b.clone()
    info = torch.full((1,), 7, dtype=torch.int32, device="cuda")  # the call clears it
    nat.check(lib.nka_cholesky_solve_f64_noalloc(_ptr(a), n, n, _ptr(x), _ptr(info), _stream()),
              "nka_cholesky_solve_f64_noalloc")
    return a, x, info


def test_no_alloc_matches_workspace_pair_on_shrinking_calls(gpu):
    # FP64 bar: 1e-10 relative; the first size has split tail tiles in the
    # workspace pair's trailing products, the smaller ones need no split
    for n, seed in ((8384, 1), (2000, 2), (700, 3)):
        print(f"case no-alloc cholesky n={n}", flush=True)
        A, b = _spd(n, seed)
        a_ref, x_ref, info_ref = _with_workspace(A, b)
        a_na, x_na, info_na = _no_alloc(A, b)
        torch.cuda.synchronize()
        assert int(info_ref.item()) == 0 and int(info_na.item()) == 0
        scale = x_ref.abs().max().item()
        assert (x_na - x_ref).abs().max().item() <= 1e-10 * scale
        L_ref, L_na = torch.tril(a_ref), torch.tril(a_na)
        assert (L_na - L_ref).abs().max().item() <= 1e-10 * L_ref.abs().max().item()
        # and the solve itself: A x = b
        assert (A @ x_na - b).abs().max().item() <= 1e-8 * b.abs().max().item()


def test_no_alloc_failure_sets_info_and_nan(gpu):
    print("case no-alloc cholesky indefinite n=300", flush=True)
    n = 300
    A, b = _spd(n, 4)
    A[150, 150] = -5.0
    _, x_na, info_na = _no_alloc(A, b)
    _, x_ref, info_ref = _with_workspace(A, b)
    torch.cuda.synchronize()
    assert int(info_na.item()) == int(info_ref.item()) > 0
    assert torch.isnan(x_na).all().item() and torch.isnan(x_ref).all().item()
    # a good call after the failure is unaffected by the scratch it left behind
    A2, b2 = _spd(500, 5)
    _, x2, info2 = _no_alloc(A2, b2)
    _, x2r, _ = _with_workspace(A2, b2)
    torch.cuda.synchronize()
    assert int(info2.item()) == 0
    assert (x2 - x2r).abs().max().item() <= 1e-10 * x2r.abs().max().item()


def test_no_alloc_rejects_too_large(gpu):
    lib = nat.lib()
    nmax = lib.nka_cholesky_solve_noalloc_max_n()
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    st = lib.nka_cholesky_solve_f64_noalloc(_ptr(t), nmax + 64, nmax + 64, _ptr(t), _ptr(info), _stream())
    assert st == -1
    assert b"exceeds the maximum" in lib.nka_last_error()
