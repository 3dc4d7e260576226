"""SR preconditioner and SGD.

* ``SR``  netket/optimizer/sr.py:62-204 with AbstractLinearPreconditioner.__call__
          (netket/optimizer/preconditioner.py:107-121): build the QGT from the
          state, solve S x = grad (default solver: CG, netket/optimizer/sr.py:121),
          keep x0/info.
* ``Sgd`` optax.sgd as wrapped by netket/optimizer/__init__.py:33-60, applied by
          apply_gradient (netket/_src/driver/abstract_optimization_driver.py:86-92):
          p <- p + (-lr * dp) in the parameter dtype.
"""

from __future__ import annotations

from . import solver as solvers
from .qgt import QGTJacobianDense


class SR:
    def __init__(self, qgt=None, solver=None, *, diag_shift=0.01, diag_scale=None,
                 solver_restart: bool = False, **kwargs):
        self.qgt_constructor = qgt if qgt is not None else QGTJacobianDense
        self.solver = solver if solver is not None else solvers.cg
        self.diag_shift = diag_shift
        self.diag_scale = diag_scale
        self.solver_restart = solver_restart
        self.qgt_kwargs = kwargs
        self.x0 = None
        self.info = None
        self._lhs = None

    def lhs_constructor(self, vstate, step=None):
        diag_shift = self.diag_shift
        if callable(diag_shift):
            if step is None:
                raise TypeError("a scheduled diag_shift needs the step argument")
            diag_shift = diag_shift(step)
        diag_scale = self.diag_scale
        if callable(diag_scale):
            if step is None:
                raise TypeError("a scheduled diag_scale needs the step argument")
            diag_scale = diag_scale(step)
        return self.qgt_constructor(vstate, diag_shift=diag_shift, diag_scale=diag_scale,
                                    **self.qgt_kwargs)

    def __call__(self, vstate, gradient, step=None):
        self._lhs = self.lhs_constructor(vstate, step)
        x0 = self.x0 if self.solver_restart else None
        self.x0, self.info = self._lhs.solve(self.solver, gradient, x0=x0)
        return self.x0

    def __repr__(self):
        return f"SR(diag_shift={self.diag_shift}, solver={getattr(self.solver, '__name__', self.solver)})"


def identity_preconditioner(vstate, gradient, step=None):
    return gradient


class Sgd:
    """Plain stochastic gradient descent: p <- p - learning_rate * dp."""

    def __init__(self, learning_rate=0.01):
        self.learning_rate = learning_rate

    def lr_at(self, step: int) -> float:
        lr = self.learning_rate
        return float(lr(step)) if callable(lr) else float(lr)

    def __repr__(self):
        return f"Sgd(learning_rate={self.learning_rate})"
