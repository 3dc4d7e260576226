"""Variational drivers.

* ``VMC``     netket/driver/vmc.py:35-161: expect_and_grad, preconditioner,
              update (netket/_src/driver/abstract_optimization_driver.py:86-92).
* ``VMC_SR``  netket/_src/driver/vmc_sr.py:86-510: SR (use_ntk=False) or
              minSR/NTK (use_ntk=True) with the linear algebra of
              netket/_src/ngd/sr_srt_common.py:44-188, sr.py:16-80, srt.py:14-96.
* ``run``     the step loop of AbstractDriver.run
              (netket/_src/driver/abstract_variational_driver.py:349-499) with
              loggers and callbacks (a callback returning False stops the run).
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import _engine as eng
from . import _native as nat
from . import _ops
from .logging import RuntimeLog
from .optimizer import solver as solvers
from .optimizer.sr import Sgd, identity_preconditioner
from .stats import Stats, statistics


def statistics_from_moments(st: Stats) -> Stats:
    """A Stats object owning a copy of ``st``'s device moments."""
    return Stats(_moments=st._moments.clone(), _shape=st._shape)


class AbstractVariationalDriver:
    def __init__(self, variational_state, optimizer, minimized_quantity_name: str = "Energy"):
        if not isinstance(optimizer, Sgd):
            raise NotImplementedError("only the Sgd optimiser is supported on the GPU path")
        self.state = variational_state
        self.optimizer = optimizer
        self._loss_name = minimized_quantity_name
        self._step_count = 0
        self._loss_stats = None
        self._dp = None

    @property
    def step_count(self) -> int:
        return self._step_count

    def reset(self):
        self.state.reset()
        self._step_count = 0

    def reset_step(self):
        self.state.reset()

    def compute_loss_and_update(self):
        raise NotImplementedError

    def resolve_fallback(self, dp):
        """The update of the last compute_loss_and_update, with the fallback
        solver's solution when the device-side nan_fallback test of
        ``cholesky_with_fallback`` fired.

        Reads the device flag: the one host synchronisation of an eager step
        that uses that solver, made after the whole step has been enqueued.
        Inside a captured step nothing is read: a flagged solution is zero
        there (no update) and the flag stays on the device in the solver's
        info."""
        pend = getattr(self, "_pending_fallback", None)
        self._pending_fallback = None
        if pend is None or torch.cuda.is_current_stream_capturing():
            return dp
        fb, redo, note = pend
        fired = bool(fb)
        note(fired)
        return redo() if fired else dp

    def _set_pending_fallback(self, info, redo, note):
        fb = info.get("solver_fallback") if isinstance(info, dict) else None
        self._pending_fallback = (fb, redo, note) if isinstance(fb, solvers.DeviceFallback) else None

    def update_parameters(self, dp_flat: torch.Tensor):
        """p <- p + (-lr * (dtype)dp) on the flat parameter buffer."""
        eng.sgd_update_(self.state.parameters_flat, dp_flat, self.optimizer.lr_at(self._step_count))
        self.state.reset()

    def _log_additional_data(self, log_data: dict):
        acc = self.state.sampler_state.acceptance
        if acc is not None:
            log_data["acceptance"] = acc

    def run(self, n_iter: int, out=None, obs=None, step_size: int = 1, show_progress: bool = False,
            callback=None, timeit: bool = False):
        if out is None:
            loggers = ()
        elif isinstance(out, (list, tuple)):
            loggers = tuple(out)
        else:
            loggers = (out,)
        callbacks = () if callback is None else (callback if isinstance(callback, (list, tuple)) else (callback,))
        for _ in range(int(n_iter)):
            self.reset_step()
            self._loss_stats, dp = self.compute_loss_and_update()
            log_data = {}
            if self._loss_stats is not None:
                log_data[self._loss_name] = self._loss_stats
            if obs:
                for name, op in obs.items():
                    log_data[name] = self.state.expect(op)
            self._log_additional_data(log_data)
            dp = self.resolve_fallback(dp)
            stop = False
            for cb in callbacks:
                if cb(self._step_count, log_data, self) is False:
                    stop = True
            if self._step_count % step_size == 0:
                for lg in loggers:
                    lg(self._step_count, log_data, self.state)
            self.update_parameters(dp)
            self._step_count += 1
            if stop:
                break
        for lg in loggers:
            if hasattr(lg, "flush"):
                lg.flush(self.state)
        return loggers

    @property
    def energy(self):
        return self._loss_stats

    # -- captured steps --------------------------------------------------------
    def _eager_step(self):
        self.reset_step()
        st, dp = self.compute_loss_and_update()
        dp = self.resolve_fallback(dp)
        self.update_parameters(dp)
        return st

    def capture_step(self, warmup: int = 2):
        """Capture one optimisation step into a HIP graph.

        ``warmup`` real steps run first on the capture stream (so every
        workspace exists), then one step is recorded; each ``replay_step`` then
        performs a full step (sampling with fresh Philox draws from the
        device-side counter, local energies, solve, update) with a single
        launch from the host.  Learning rate and diag_shift are fixed at
        capture time.
        """
        self._graph_stream = torch.cuda.Stream()
        self._graph_stream.wait_stream(torch.cuda.current_stream())
        # scratch buffers owned by this driver for as long as the graph lives
        # (stream handles come from a pool and are shared between drivers)
        self._graph_ws = _ops.Workspace()
        self.warmup_stats = []
        with _ops.workspace_scope(self._graph_ws):
            with torch.cuda.stream(self._graph_stream):
                for _ in range(warmup):
                    self.warmup_stats.append(self._eager_step())
                    self._step_count += 1
            torch.cuda.current_stream().wait_stream(self._graph_stream)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=self._graph_stream):
                st = self._eager_step()
        self._graph = g
        self._graph_stats = st
        return g

    def replay_step(self):
        """One captured step; returns its Stats (resolved lazily)."""
        self._graph.replay()
        self._step_count += 1
        st = self._graph_stats
        return statistics_from_moments(st)


class VMC(AbstractVariationalDriver):
    """Energy minimisation by VMC with an optional preconditioner (e.g. SR)."""

    def __init__(self, hamiltonian, optimizer, *, variational_state, preconditioner=None):
        super().__init__(variational_state, optimizer)
        self._ham = hamiltonian
        self.preconditioner = preconditioner if preconditioner is not None else identity_preconditioner

    def compute_loss_and_update(self):
        self.state.reset()
        st, g, _ = self.state.expect_and_forces_flat(self._ham)
        if self.preconditioner is identity_preconditioner:
            dp = g
        else:
            dp = self.preconditioner(self.state, g, self._step_count)
            if isinstance(dp, dict):
                dp = self.state.flatten_tree(dp)
            pc = self.preconditioner
            info = getattr(pc, "info", None)

            def redo():
                x = info["solver_fallback"].solve()
                pc.x0 = x
                return self.state.flatten_tree(x) if isinstance(x, dict) else x

            def note(fired):
                pc.info = {**info, "solver_fallback": fired}

            self._set_pending_fallback(info, redo, note)
        return st, dp


class VMC_SR(AbstractVariationalDriver):
    """VMC with SR or minSR (NTK) natural-gradient updates."""

    def __init__(self, hamiltonian, optimizer, *, diag_shift, variational_state, proj_reg=None,
                 momentum=None, linear_solver=solvers.cholesky_with_fallback, linear_solver_fn=None,
                 chunk_size_bwd=None, mode=None, use_ntk=None, on_the_fly=None,
                 ntk_method="structured", qgt_method="structured"):
        super().__init__(variational_state, optimizer)
        if ntk_method not in ("structured", "jacobian"):
            raise ValueError("ntk_method must be 'structured' or 'jacobian'")
        if qgt_method not in ("structured", "jacobian"):
            raise ValueError("qgt_method must be 'structured' or 'jacobian'")
        if linear_solver_fn is not None:
            linear_solver = linear_solver_fn
        if mode not in (None, "real"):
            raise NotImplementedError("only mode='real' is supported")
        if use_ntk is None:
            use_ntk = variational_state.n_parameters > variational_state.n_samples
        if on_the_fly is None:
            on_the_fly = False
        elif on_the_fly and not use_ntk:
            raise ValueError("`on_the_fly=True` is only supported when `use_ntk=True`.")
        if on_the_fly:
            # T is formed from the model's layer structure without storing the
            # Jacobian (netket/_src/ngd/srt_onthefly.py:31 computes the same T
            # with vjp/jvp products)
            if ntk_method != "structured":
                raise ValueError("on_the_fly=True needs ntk_method='structured'")
        self._ham = hamiltonian
        self.diag_shift = diag_shift
        self.proj_reg = proj_reg
        self.momentum = momentum
        self.use_ntk = bool(use_ntk)
        self.on_the_fly = bool(on_the_fly)
        self.linear_solver = linear_solver
        # NTK T from the RBM's structure (FP64, inner sizes n_hidden and n_visible)
        # or from the explicit Jacobian (bf16x3 / FP64 Gram over n_params)
        self.ntk_method = ntk_method
        # S of an FP32 RBM from its spin-pair x activation-pair structure, or
        # from the explicit Jacobian (bf16x3 / FP64 Gram over n_samples)
        self.qgt_method = qgt_method
        self._old_updates = None
        self._jac = None
        self._side = None  # second stream of the structured S
        self.info = None
        # when a dict, each step leaves its local energies, right-hand side,
        # solution, update and linear system there (for error reports)
        self.record = None

    def _hyper(self, v):
        return v(self._step_count) if callable(v) else v

    def compute_loss_and_update(self, eloc=None):
        """(Stats, update).  ``eloc`` (n_chains, chain_length) may be given when
        the local energies of the current samples were computed elsewhere (the
        serving loop's shared launches); otherwise they are computed here."""
        vs = self.state
        diag_shift = float(self._hyper(self.diag_shift))
        proj_reg = self._hyper(self.proj_reg)
        momentum = self._hyper(self.momentum)
        model, flat = vs.model, vs.parameters_flat
        structured_s = (not self.use_ntk and self.qgt_method == "structured"
                        and eng.qgt_structured_available(model, flat))
        pc = vs._prepared()
        sig, t = pc["sigma"], pc["t"]
        ns = sig.shape[0]
        need_tT = not self.use_ntk
        tT = eng.prepare(model, flat, sig, row_major=False, transposed=True)[1] if need_tT else None
        S_pre = None
        if structured_s and eloc is None:
            # S depends only on the samples: it is formed on a second stream
            # while this stream computes the local energies, the statistics and
            # the contractions (matrix cores against the scalar unit and
            # latency-bound work), and joined before the solve
            main = torch.cuda.current_stream()
            if self._side is None:
                self._side = torch.cuda.Stream()
            np_ = model.layout(sig.shape[1])["n_params"]
            S_pre = torch.empty((np_, np_), dtype=torch.float64, device=sig.device)
            self._side.wait_stream(main)
            with torch.cuda.stream(self._side):
                eng.qgt_structured(model, flat, sig, tT, 1.0 / ns, diag_shift, out=S_pre)
            # the allocator keeps these blocks until the second stream's work on
            # them is done, also if this step stops before the join
            for x in (S_pre, sig, tT, flat):
                x.record_stream(self._side)
        if eloc is None:
            eloc = vs.local_estimators(self._ham)
        st = statistics(eloc)
        e = eloc.reshape(-1)
        ones = torch.full((1, ns), 1.0 / ns, dtype=torch.float64, device=e.device)
        mean = eng.contract(model, flat, sig, t, ones)[0]
        # dv = 2 (E - <E>) / sqrt(Ns)   (sr_srt_common.py:75-77)
        dv = eng.centre_scale(e, st.mean_device, 2.0 / math.sqrt(ns))
        # contraction weights dv / sqrt(Ns) = 2 (E - <E>) / Ns
        w_dv = eng.centre_scale(e, st.mean_device, 2.0 / ns)
        structured = self.use_ntk and self.ntk_method == "structured"
        if not structured and not structured_s:
            kind = nat.OUT_BF16X2 if flat.dtype == torch.float32 else nat.OUT_F64
            layout = nat.LAYOUT_SP if self.use_ntk else nat.LAYOUT_PS
            self._jac = eng.jacobian(model, flat, sig, t, tT, mean, 1.0 / math.sqrt(ns), layout,
                                     kind, self._jac)
        X = self._jac
        if momentum is not None:
            # the previous update, kept so that a fallback solve can redo this step's momentum terms
            self._momentum_buffer(flat)
            self._old_saved = self._old_updates.clone()
        if not self.use_ntk:
            if proj_reg is not None:
                raise ValueError("proj_reg not implemented for SR")
            # F = O_L^T dv  (the contraction weights are dv / sqrt(Ns))
            F = eng.contract(model, flat, sig, t, w_dv.reshape(1, -1), mean)[0]
            if momentum is not None:
                # F + eps mu old_updates (netket/_src/ngd/sr.py:45-47)
                eng.axpy_(F, diag_shift * float(momentum), self._old_updates)
            if structured_s:
                build_s = lambda: eng.qgt_structured(model, flat, sig, tT, 1.0 / ns, diag_shift)
                if S_pre is not None:
                    torch.cuda.current_stream().wait_stream(self._side)
                system = solvers.DenseSystem(S_pre, builder=build_s) if S_pre is not None else \
                    solvers.DenseSystem(builder=build_s)
            else:
                system = solvers.DenseSystem(builder=lambda: eng.gram(X, diag_shift))
            rhs = F

            def finish(x):
                if momentum is not None:
                    self._old_updates.copy_(x)  # old_updates = updates (sr.py:67-68)
                return x
        else:
            if momentum is not None:
                # dv -= mu O_L old_updates (netket/_src/ngd/srt.py:34-35)
                eng.jvp(model, flat, sig, t, self._old_updates, mean,
                        -float(momentum) / math.sqrt(ns), 1.0, dv)
            offset = 0.0 if proj_reg is None else float(proj_reg) / ns
            if structured:
                system = solvers.DenseSystem(builder=lambda: eng.ntk_structured(
                    model, flat, sig, t, mean, 1.0 / ns, diag_shift, offset))
            else:
                system = solvers.DenseSystem(builder=lambda: eng.gram(X, diag_shift, offset))
            rhs = dv

            def finish(y):
                # updates = O_L^T y (srt.py:79), from the analytic Jacobian
                dp = eng.contract(model, flat, sig, t, eng.scale(y, 1.0 / math.sqrt(ns)).reshape(1, -1),
                                  mean)[0]
                if momentum is not None:
                    # updates += mu old_updates; old_updates = updates (srt.py:80-82)
                    eng.axpy_(dp, float(momentum), self._old_updates)
                    self._old_updates.copy_(dp)
                return dp
        y, info = self._solve(system, rhs)
        dp = finish(y)
        self.info = info
        if self.record is not None:
            self.record.update(eloc=e, rhs=rhs, solution=y, dp=dp, system=system)

        def redo():
            if momentum is not None:
                self._old_updates.copy_(self._old_saved)
            return finish(info["solver_fallback"].solve())

        def note(fired):
            self.info = {**info, "solver_fallback": fired}

        self._set_pending_fallback(info, redo, note)
        return st, dp

    def _momentum_buffer(self, flat):
        """old_updates, zero before the first step (sr_srt_common.py:168-169);
        one persistent buffer, so a captured step reads the previous replay's."""
        if self._old_updates is None or self._old_updates.numel() != flat.numel():
            self._old_updates = torch.zeros(flat.numel(), dtype=torch.float64, device=flat.device)
        return self._old_updates

    def _solve(self, system, b):
        res = self.linear_solver(system, b)
        if isinstance(res, tuple):
            x, info = res
            return x, (info or {})
        return res, {}
