"""Optimisers and the SR preconditioner (netket/optimizer/)."""

from __future__ import annotations

from . import solver
from .qgt import QGTAuto, QGTJacobianDense, QGTJacobianDenseT
from .sr import SR, Sgd, identity_preconditioner

__all__ = ["solver", "QGTAuto", "QGTJacobianDense", "QGTJacobianDenseT", "SR", "Sgd",
           "identity_preconditioner"]
