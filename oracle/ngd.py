"""SR / minSR linear algebra — oracle (test infrastructure only).

* ``prepare``    _prepare_input (netket/_src/ngd/sr_srt_common.py:44-90), real mode.
* ``sr_update``  _compute_sr_update (netket/_src/ngd/sr.py:16-80):
                 solve (O_L^T O_L + eps I) x = O_L^T dv.
* ``srt_update`` _compute_srt_update (netket/_src/ngd/srt.py:14-96):
                 x = O_L^T (O_L O_L^T + eps I [+ proj_reg/n])^{-1} dv.
* ``qgt_dense``  QGTJacobianDenseT.to_dense (netket/optimizer/qgt/qgt_jacobian_dense.py:136-161).
* ``cholesky_solve`` cho_factor/cho_solve (netket/optimizer/solver/solvers.py:241-270), scipy LAPACK.
* ``pinv_smooth`` netket/optimizer/solver/solvers.py:32-137.
* ``cg``         jax.scipy.sparse.linalg.cg semantics (tol relative to ||b||,
                 atol, maxiter = 10 n, x0 = 0), the default SR solver
                 (netket/optimizer/sr.py:121).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla


def _weights(n: int, weights):
    return np.full(n, 1.0 / n) if weights is None else np.asarray(weights, dtype=np.float64)


def centred_jacobian(O: np.ndarray, weights=None) -> np.ndarray:
    """sqrt(w_s) (O_s - sum_t w_t O_t): with w = 1/Ns this is
    (O - mean over samples) / sqrt(Ns)  (netket/jax/_jacobian/logic.py:365-373)."""
    w = _weights(O.shape[0], weights)
    return np.sqrt(w)[:, None] * (O - (w @ O)[None, :])


def prepare(O: np.ndarray, eloc: np.ndarray, weights=None):
    """(O_L, dv) of _prepare_input (netket/_src/ngd/sr_srt_common.py:44-90):
    dv = 2 sqrt(w) (E - sum_s w_s E_s), i.e. 2 (E - <E>) / sqrt(Ns) for the
    uniform sample weights.  O_L^T dv is then the energy gradient of
    ``rbm.forces_to_grad``; with the exact weights |psi|^2/Z over the full
    basis it is the exact gradient (pinned by finite differences in the
    oracle tests)."""
    eloc = np.asarray(eloc, dtype=np.float64)
    w = _weights(eloc.shape[0], weights)
    OL = centred_jacobian(O, w)
    dv = 2.0 * np.sqrt(w) * (eloc - np.sum(w * eloc))
    return OL, dv


def cholesky_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    c, low = sla.cho_factor(A, lower=False)
    return sla.cho_solve((c, low), b)


def sr_update(OL, dv, diag_shift, momentum=None, old_updates=None):
    """SR update; with SPRING momentum F += eps mu old_updates (sr.py:43-47),
    and the new old_updates is the returned x (sr.py:67-68)."""
    F = OL.T @ dv
    if momentum is not None:
        F = F + diag_shift * momentum * old_updates
    S = OL.T @ OL + diag_shift * np.eye(OL.shape[1])
    return cholesky_solve(S, F), S, F


def srt_update(OL, dv, diag_shift, proj_reg=None, momentum=None, old_updates=None):
    """minSR update; with SPRING momentum dv -= mu O_L old_updates before the
    solve and updates += mu old_updates after it (srt.py:34-35, 80-82)."""
    if momentum is not None:
        dv = dv - momentum * (OL @ old_updates)
    T = OL @ OL.T + diag_shift * np.eye(OL.shape[0])
    if proj_reg is not None:
        T = T + proj_reg / T.shape[0]
    y = cholesky_solve(T, dv)
    upd = OL.T @ y
    if momentum is not None:
        upd = upd + momentum * old_updates
    return upd, T, y


def qgt_dense(OL, diag_shift):
    return OL.T @ OL + diag_shift * np.eye(OL.shape[1])


def pinv_smooth(A, b, rtol=1e-14, rtol_smooth=1e-14):
    w, U = np.linalg.eigh(A)
    w_inv = np.where(np.abs(w / w[-1]) > rtol, 1.0 / w, 0.0)
    reg = 1.0 / (1.0 + (rtol_smooth / np.abs(w / w[-1])) ** 6)
    return U @ (w_inv * reg * (U.T @ b))


def cg(matvec, b, x0=None, tol=1e-5, atol=0.0, maxiter=None):
    b = np.asarray(b, dtype=np.float64)
    n = b.size
    if maxiter is None:
        maxiter = 10 * n
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - matvec(x)
    p = r.copy()
    gamma = r @ r
    atol2 = max(tol**2 * (b @ b), atol**2)
    k = 0
    while gamma > atol2 and k < maxiter:
        Ap = matvec(p)
        alpha = gamma / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        gamma_new = r @ r
        beta = gamma_new / gamma
        p = r + beta * p
        gamma = gamma_new
        k += 1
    return x, k
