"""Dense Jacobian quantum geometric tensor (QGTJacobianDense).

Follows QGTJacobian_DefaultConstructor and QGTJacobianDenseT
(netket/optimizer/qgt/qgt_jacobian.py:30-206,
netket/optimizer/qgt/qgt_jacobian_dense.py:29-161, 251-270) for real
parameters ("real" mode): O_L = (O - <O>) / sqrt(Ns) with FP64 centring,
S = O_L^T O_L + diag_shift I.

For FP32 RBMs S is formed from the model's structure without a Jacobian: each
log-derivative is (augmented spin) x (augmented tanh), so O^T O is unpacked
from the product of the spin-pair and activation-pair tables over the samples
(nka_rbm_qgt, bf16 matrix cores, FP64 centring).  Other FP32 models store O_L
parameter-major (X = O_L^T, Np x Ns) as a bf16 hi/lo split and form S with
three BF16 MFMA products per tile (FP32 accumulate); FP64 models keep O_L and
S in FP64 (FP64 MFMA).  S is built on demand and not kept after a Cholesky
solve consumed it.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _engine as eng
from .. import _native as nat
from .. import _ops


class LinearOperator:
    def __call__(self, vec):
        return self @ vec

    def solve(self, solve_fun, y, *, x0=None, **kwargs):
        """Solve S x = y with ``solve_fun(S, y, x0=...) -> (x, info)``
        (netket/optimizer/linear_operator.py:86-103)."""
        return solve_fun(self, y, x0=x0, **kwargs)


class QGTJacobianDenseT(LinearOperator):
    def __init__(self, vstate, *, diag_shift: float = 0.0, diag_scale=None, mode=None,
                 holomorphic=None, chunk_size=None):
        if mode not in (None, "real"):
            raise NotImplementedError("only real-parameter ('real' mode) QGTs are supported")
        self.vstate = vstate
        # to_shift_offset (netket/optimizer/qgt/qgt_jacobian_common.py:21-34): with a
        # nonzero diag_scale the solved system is D^-1 S D^-1 + diag_scale I with
        # D_kk = sqrt(S_kk + diag_shift / diag_scale)
        diag_shift = 0.0 if diag_shift is None else float(diag_shift)
        if diag_scale is None or float(diag_scale) == 0.0:
            self.diag_shift, self._offset = diag_shift, None
        else:
            self.diag_shift, self._offset = float(diag_scale), diag_shift / float(diag_scale)
        self.scale = None
        s = vstate.samples
        C, L, N = s.shape
        self.n_samples = C * L
        sig = s.reshape(C * L, N)
        model, flat = vstate.model, vstate.parameters_flat
        t, tT = eng.prepare(model, flat, sig, row_major=True, transposed=True)
        self.X = None
        if eng.qgt_structured_available(model, flat):
            self._structured = (model, flat, sig, tT)
            self.n_params = model.layout(N)["n_params"]
        else:
            self._structured = None
            ones = torch.full((1, self.n_samples), 1.0 / self.n_samples, dtype=torch.float64,
                              device=sig.device)
            self.mean = eng.contract(model, flat, sig, t, ones)[0]
            kind = nat.OUT_BF16X2 if flat.dtype == torch.float32 else nat.OUT_F64
            self.X = eng.jacobian(model, flat, sig, t, tT, self.mean, eng.inv_sqrt(self.n_samples),
                                  nat.LAYOUT_PS, kind)
            self.n_params = self.X.rows
        self._dense = None

    def _gram(self, shift: float) -> torch.Tensor:
        """S + shift I on the lower triangle (FP64)."""
        if self._structured is not None:
            model, flat, sig, tT = self._structured
            return eng.qgt_structured(model, flat, sig, tT, 1.0 / self.n_samples, shift)
        return eng.gram(self.X, shift)

    # -- matrices ---------------------------------------------------------------
    def lower(self) -> torch.Tensor:
        """A fresh FP64 matrix whose lower triangle is the system that is
        solved: S + diag_shift I, or D^-1 S D^-1 + diag_shift I with a
        diag_scale (which also sets ``self.scale`` = diag(D))."""
        if self._offset is None:
            return self._gram(self.diag_shift)
        a = self._gram(0.0)
        self.scale = torch.empty(a.shape[0], dtype=torch.float64, device=a.device)
        _ops.qgt_rescale_(a, self._offset, self.diag_shift, 0, self.scale)
        return a

    def to_dense(self) -> torch.Tensor:
        """S + diag_shift I, or S + diag_shift D^2 with a diag_scale
        (qgt_jacobian_dense.py:136-161)."""
        if self._dense is None:
            if self._offset is None:
                a = self.lower()
            else:
                a = self._gram(0.0)
                self.scale = torch.empty(a.shape[0], dtype=torch.float64, device=a.device)
                _ops.qgt_rescale_(a, self._offset, self.diag_shift, 1, self.scale)
            _ops.symmetrize_(a)
            self._dense = a
        return self._dense

    def solve(self, solve_fun, y, *, x0=None, **kwargs):
        """_solve of netket/optimizer/qgt/qgt_jacobian_dense.py:93-133: with a
        diag_scale the solver sees the rescaled system, y / D and x0 D, and the
        solution is divided by D."""
        if self._offset is None:
            return solve_fun(self, y, x0=x0, **kwargs)
        from .solver import DenseSystem

        vs = self.vstate
        yf = vs.flatten_tree(y) if isinstance(y, dict) else y.to(torch.float64).contiguous()
        a = self.lower()
        sys_ = DenseSystem(a, builder=self.lower)
        ys = _ops.mul_div(yf, self.scale, divide=True)
        x0s = None
        if x0 is not None:
            x0f = vs.flatten_tree(x0) if isinstance(x0, dict) else x0.to(torch.float64).contiguous()
            x0s = _ops.mul_div(x0f, self.scale, divide=False)
        out, info = solve_fun(sys_, ys, x0=x0s, **kwargs)
        out = _ops.mul_div(out, self.scale, divide=True)
        if isinstance(y, dict):
            out = vs.unflatten(out)
        return out, info

    def __matmul__(self, vec):
        v = self.vstate.flatten_tree(vec) if not isinstance(vec, torch.Tensor) else vec.to(torch.float64)
        out = _ops.gemv(self.to_dense(), v.contiguous())
        if isinstance(vec, dict):
            return self.vstate.unflatten(out)
        return out

    def __repr__(self):
        if self._offset is not None:
            return f"QGTJacobianDense(diag_scale={self.diag_shift}, offset={self._offset}, mode=real)"
        return f"QGTJacobianDense(diag_shift={self.diag_shift}, mode=real)"


def QGTJacobianDense(vstate=None, *, diag_shift: float = 0.0, diag_scale=None, mode=None,
                     holomorphic=None, chunk_size=None, **kwargs):
    """Constructor (or partial when called without a state)."""
    if vstate is None:
        def ctor(vs, **kw):
            args = dict(diag_shift=diag_shift, diag_scale=diag_scale, mode=mode,
                        holomorphic=holomorphic, chunk_size=chunk_size)
            args.update(kw)
            return QGTJacobianDenseT(vs, **args)
        return ctor
    return QGTJacobianDenseT(vstate, diag_shift=diag_shift, diag_scale=diag_scale, mode=mode,
                             holomorphic=holomorphic, chunk_size=chunk_size)


QGTAuto = QGTJacobianDense
