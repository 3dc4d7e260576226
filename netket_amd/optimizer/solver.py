"""Dense solvers with the reference signature ``solver(A, b, *, x0=None) -> (x, info)``
(``SolverT``, netket/optimizer/linear_operator.py:32; netket/optimizer/solver/solvers.py).

``A`` is a QGT linear operator or a dense FP64 device matrix (for VMC_SR, whose
matrix is formed explicitly); ``b`` a flat FP64 device vector or a
parameter dict.  All factorisations and iterations run on the GPU.
"""

from __future__ import annotations

import torch

from .. import _native as nat
from .. import _ops


class DenseSystem:
    """A dense SPD matrix (lower triangle valid) for the solvers.

    ``builder()`` returns a fresh FP64 matrix whose lower triangle holds the
    system; a Cholesky solve factors such a matrix in place, so every
    consumer asks for its own copy (rebuilt from the Jacobian rather than
    copied when a builder is given).
    """

    def __init__(self, lower: torch.Tensor | None = None, builder=None):
        self._first = lower
        self._builder = builder
        self._full = None

    def lower(self) -> torch.Tensor:
        if self._first is not None:
            a, self._first = self._first, None
            if self._builder is None:
                self._keep = a.clone()
            return a
        if self._builder is not None:
            return self._builder()
        return self._keep.clone()

    def to_dense(self) -> torch.Tensor:
        if self._full is None:
            a = self.lower()
            _ops.symmetrize_(a)
            self._full = a
        return self._full


def _system(A, b):
    """(object with .lower()/.to_dense(), flat FP64 b, unflatten)."""
    from .qgt import QGTJacobianDenseT

    if isinstance(A, QGTJacobianDenseT):
        vs = A.vstate
        if isinstance(b, dict):
            bf = vs.flatten_tree(b)
            return A, bf.contiguous(), vs.unflatten
        return A, b.to(torch.float64).contiguous(), (lambda x: x)
    if isinstance(A, DenseSystem):
        return A, b.to(torch.float64).contiguous(), (lambda x: x)
    if isinstance(A, torch.Tensor):
        return DenseSystem(A.to(torch.float64).contiguous().clone()), b.to(torch.float64).contiguous(), (lambda x: x)
    raise TypeError(f"unsupported linear operator {type(A).__name__}")


def _chol_solve(sys_, b: torch.Tensor):
    a = sys_.lower()
    n = a.shape[0]
    info = torch.zeros(1, dtype=torch.int32, device=a.device)
    ws = _ops.potrf_(a, info)
    x = b.clone()
    _ops.potrs_(a, ws, x, info)
    return x, info


def cholesky(A, b, *, lower: bool = False, x0=None):
    """Cholesky factorisation and solve (netket/optimizer/solver/solvers.py:241-270)."""
    sys_, bf, unflat = _system(A, b)
    x, _ = _chol_solve(sys_, bf)
    return unflat(x), None


def pinv_smooth(A, b, *, rtol: float = 1e-14, rtol_smooth: float = 1e-14, x0=None,
                rcond=None, rcond_smooth=None):
    """Smoothed pseudo-inverse from the eigendecomposition
    (netket/optimizer/solver/solvers.py:32-137).

    The eigendecomposition is ``torch.linalg.eigh`` (rocSOLVER), not a kernel
    of this library: it is the rarely taken fallback of
    ``cholesky_with_fallback`` and of explicit calls, never part of the
    timed step.
    """
    if rcond is not None:
        rtol = rcond
    if rcond_smooth is not None:
        rtol_smooth = rcond_smooth
    sys_, bf, unflat = _system(A, b)
    a = sys_.to_dense()
    w, U = torch.linalg.eigh(a)
    wmax = w[-1]
    rel = torch.abs(w / wmax)
    w_inv = torch.where(rel > rtol, 1.0 / w, torch.zeros_like(w))
    reg = 1.0 / (1.0 + (rtol_smooth / rel) ** 6)
    x = U @ ((w_inv * reg) * (U.T @ bf))
    info = {"eval_min": w[0], "eval_max": wmax, "rank": torch.sum(w > rtol * wmax),
            "cond_number": wmax / w[0]}
    return unflat(x), info


class DeviceFallback:
    """The outcome of the nan_fallback test, decided on the device.

    ``flag`` is a device int32 (nonzero: the right-hand side or the primary
    solution held NaN/Inf, or the factorisation failed).  The primary
    solution has already been replaced by zeros on the device in that case,
    so nothing downstream consumes NaN; ``solve()`` runs the fallback solver
    on the same system when the caller finds the flag set.  Reading the flag
    is the caller's choice of synchronisation point (the reference decides
    with ``jax.lax.cond`` on the device, netket/_src/solvers/nan_fallback.py:34-38).
    """

    def __init__(self, flag: torch.Tensor, fallback):
        self.flag = flag
        self._fallback = fallback

    def __bool__(self) -> bool:
        return bool(self.flag.item() != 0)

    def solve(self):
        return self._fallback()


def cholesky_with_fallback(A, b, *, rtol: float = 1e-14, rtol_smooth: float = 1e-14, x0=None):
    """Cholesky, falling back to pinv_smooth on NaN/Inf
    (netket/optimizer/solver/solvers.py:304-340, netket/_src/solvers/nan_fallback.py:22-45).

    Nothing here synchronises with the host, so the call can be captured in a
    HIP graph.  The test is made on the device: a non-finite b, a failed
    factorisation (the solve then returns NaN) or a non-finite x set the flag,
    and a flagged x is zeroed on the device.  ``info["solver_fallback"]`` is a
    ``DeviceFallback``: ``bool()`` of it reads the flag, and its ``solve()``
    returns the pinv_smooth solution of the same system.  The drivers resolve
    it at the end of each eager step.
    """
    sys_, bf, unflat = _system(A, b)
    flag = torch.zeros(1, dtype=torch.int32, device=bf.device)
    _ops.check_finite(bf, flag)
    x, info = _chol_solve(sys_, bf)
    _ops.check_finite(x, flag)
    _ops.select_on_flag(flag, None, x)
    fb = DeviceFallback(flag, lambda: pinv_smooth(sys_, bf, rtol=rtol, rtol_smooth=rtol_smooth)[0])
    return unflat(x), {"solver_fallback": fb}


def solve(A, b, *, x0=None):
    """Dense solve of the SPD system (via Cholesky)."""
    return cholesky(A, b, x0=x0)


def cg(A, b, x0=None, *, tol: float = 1e-5, atol: float = 0.0, maxiter: int | None = None, M=None,
       check_every: int = 8):
    """Conjugate gradient on the device (jax.scipy.sparse.linalg.cg semantics),
    from x0 = 0 or from a given x0 (the previous solution that SR passes with
    solver_restart=True, netket/optimizer/preconditioner.py:118)."""
    if M is not None:
        raise NotImplementedError("preconditioned CG is not supported")
    sys_, bf, unflat = _system(A, b)
    a = sys_.to_dense()
    n = bf.numel()
    if maxiter is None:
        maxiter = 10 * n
    x = torch.empty_like(bf)
    ws = _ops.WORKSPACE.get("cg", nat.lib().nka_cg_workspace(n))
    status = torch.empty(8, dtype=torch.float64, device=bf.device)
    if x0 is None:
        nat.call("nka_cg_start", _ops.ptr(bf), n, float(tol), float(atol), int(maxiter), _ops.ptr(x),
                 _ops.ptr(ws), _ops.stream())
    else:
        if isinstance(x0, dict):
            x0 = A.vstate.flatten_tree(x0)
        x0f = x0.to(torch.float64).contiguous()
        nat.call("nka_cg_start_x0", _ops.ptr(a), a.stride(0), _ops.ptr(bf), n, float(tol), float(atol),
                 int(maxiter), _ops.ptr(x0f), _ops.ptr(x), _ops.ptr(ws), _ops.stream())
    done = 0
    while True:
        nat.call("nka_cg_status", _ops.ptr(ws), n, _ops.ptr(status), _ops.stream())
        st = status.cpu()
        if st[6] != 0.0:
            break
        nat.call("nka_cg_iterate", _ops.ptr(a), a.stride(0), n, _ops.ptr(x), _ops.ptr(ws),
                 check_every, _ops.stream())
        done += check_every
    return unflat(x), {"iterations": int(st[5]), "residual2": float(st[0])}


jax_cg = cg
