"""Ansatze of the VMC path: the RBM (netket/models/rbm.py:32-81) and the
two-layer dense network of the J1-J2 workload.

``log psi(s) = sum_j log_cosh(b_j + sum_i s_i W_ij) + sum_i a_i s_i`` with
``log_cosh(x) = |x| + log1p(exp(-2|x|)) - log 2`` (netket/nn/activation.py:78-84).

Parameters are stored in one flat device buffer in the order ``jax``'s
``ravel_pytree`` gives the flax parameter dict (keys sorted):
``Dense/bias`` (M), ``Dense/kernel`` (N, M), ``visible_bias`` (N).  The Jacobian
columns and the SR update use the same order, so an update is applied to the
buffer directly.  ``parameters`` exposes the flax-shaped dict as views.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import _ops


class RBM:
    """Restricted Boltzmann machine with log-cosh hidden units."""

    kind = "rbm"
    metropolis_symbol = "nka_rbm_metropolis"

    def __init__(self, alpha: float = 1, param_dtype=np.float64, use_hidden_bias: bool = True,
                 use_visible_bias: bool = True, kernel_init_std: float = 0.01,
                 hidden_bias_init_std: float = 0.01, visible_bias_init_std: float = 0.01):
        self.alpha = alpha
        self.param_dtype = np.dtype(param_dtype)
        if self.param_dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise NotImplementedError("RBM parameters must be float32 or float64 (real)")
        self.use_hidden_bias = bool(use_hidden_bias)
        self.use_visible_bias = bool(use_visible_bias)
        self.kernel_init_std = float(kernel_init_std)
        self.hidden_bias_init_std = float(hidden_bias_init_std)
        self.visible_bias_init_std = float(visible_bias_init_std)

    # -- layout --------------------------------------------------------------
    def n_hidden(self, N: int) -> int:
        return int(self.alpha * N)

    def layout(self, N: int) -> dict:
        """Offsets of each parameter in the flat buffer (ravel order)."""
        M = self.n_hidden(N)
        off = 0
        lay = {"N": N, "M": M}
        if self.use_hidden_bias:
            lay["bias"] = (off, (M,))
            off += M
        lay["kernel"] = (off, (N, M))
        off += N * M
        if self.use_visible_bias:
            lay["visible_bias"] = (off, (N,))
            off += N
        lay["n_params"] = off
        return lay

    def torch_dtype(self) -> torch.dtype:
        return _ops.torch_dtype(self.param_dtype)

    def init(self, seed: int, N: int) -> torch.Tensor:
        """Flat parameter buffer with entries drawn from N(0, std^2).

        Each parameter block is drawn from its own numpy Philox stream keyed by
        (seed, block index), in ravel order.
        """
        lay = self.layout(N)
        flat = np.empty(lay["n_params"], dtype=self.param_dtype)
        stds = {"bias": self.hidden_bias_init_std, "kernel": self.kernel_init_std,
                "visible_bias": self.visible_bias_init_std}
        for k, name in enumerate(("bias", "kernel", "visible_bias")):
            if name in lay:
                off, shape = lay[name]
                n = int(np.prod(shape))
                rng = np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), k]))
                flat[off : off + n] = rng.normal(0.0, stds[name], size=n).astype(self.param_dtype)
        return torch.as_tensor(flat, device=_ops.device())

    def views(self, flat: torch.Tensor, N: int) -> dict:
        lay = self.layout(N)
        out = {}
        if "bias" in lay:
            o, s = lay["bias"]
            out.setdefault("Dense", {})["bias"] = flat[o : o + s[0]]
        o, s = lay["kernel"]
        out.setdefault("Dense", {})["kernel"] = flat[o : o + s[0] * s[1]].view(s)
        if "visible_bias" in lay:
            o, s = lay["visible_bias"]
            out["visible_bias"] = flat[o : o + s[0]]
        return out

    def flatten(self, params: dict, N: int) -> torch.Tensor:
        """Flat buffer from a flax-shaped parameter dict."""
        lay = self.layout(N)
        flat = torch.empty(lay["n_params"], dtype=self.torch_dtype(), device=_ops.device())
        for name, path in (("bias", ("Dense", "bias")), ("kernel", ("Dense", "kernel")),
                           ("visible_bias", ("visible_bias",))):
            if name not in lay:
                continue
            v = params
            for p in path:
                v = v[p]
            o, s = lay[name]
            flat[o : o + int(np.prod(s))] = torch.as_tensor(np.asarray(v) if not isinstance(v, torch.Tensor) else v,
                                                            device=flat.device).reshape(-1).to(flat.dtype)
        return flat

    def native(self, flat: torch.Tensor, N: int) -> nat.NkaRbm:
        lay = self.layout(N)
        if flat.dtype != self.torch_dtype() or not flat.is_contiguous():
            raise ValueError("parameter buffer has the wrong dtype or is not contiguous")
        m = nat.NkaRbm()
        m.n_visible = N
        m.n_hidden = lay["M"]
        m.dtype = _ops.dtype_code(flat.dtype)
        m.params = flat.data_ptr()
        m.bias_off = lay["bias"][0] if "bias" in lay else -1
        m.kernel_off = lay["kernel"][0]
        m.visible_off = lay["visible_bias"][0] if "visible_bias" in lay else -1
        return m

    # -- evaluation ---------------------------------------------------------------
    def apply(self, flat: torch.Tensor, sigma) -> torch.Tensor:
        """log psi of a batch of configurations (..., N) on the GPU."""
        N = sigma.shape[-1]
        xs = _ops.as_states(sigma, N)
        batch = xs.shape[:-1]
        xf = xs.reshape(-1, N)
        out = torch.empty(xf.shape[0], dtype=flat.dtype, device=xf.device)
        m = self.native(flat, N)
        nat.call("nka_rbm_log_psi", _ops.C.byref(m), _ops.ptr(xf), xf.shape[0], _ops.ptr(out),
                 _ops.stream())
        return out.reshape(batch)

    def __repr__(self):
        return f"RBM(alpha={self.alpha}, param_dtype={self.param_dtype})"


class FFNN:
    """Two-layer dense network ``Dense(H1) -> log_cosh -> Dense(H2) -> log_cosh -> sum``.

    The user flax module of the J1-J2 workload (the pattern of
    docs/tutorials/gs-heisenberg.ipynb "Model2"), with ``nn.Dense`` layers
    (netket/nn/blocks/mlp.py:105-115) and log_cosh
    (netket/nn/activation.py:78-84).  ``hidden_dims`` are the widths H1, H2
    of the two layers.  Flat parameters in ravel order: Dense_0/bias (H1),
    Dense_0/kernel (N, H1), Dense_1/bias (H2), Dense_1/kernel (H1, H2).
    """

    kind = "ffnn"
    metropolis_symbol = "nka_ffnn_metropolis"

    def __init__(self, hidden_dims=(400, 400), param_dtype=np.float32, use_bias: bool = True,
                 kernel_init_std: float = 0.01, bias_init_std: float = 0.01):
        if len(hidden_dims) != 2:
            raise NotImplementedError("the GPU kernels implement two hidden layers")
        self.hidden_dims = tuple(int(h) for h in hidden_dims)
        self.param_dtype = np.dtype(param_dtype)
        if self.param_dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise NotImplementedError("FFNN parameters must be float32 or float64 (real)")
        self.use_bias = bool(use_bias)
        self.kernel_init_std = float(kernel_init_std)
        self.bias_init_std = float(bias_init_std)

    def layout(self, N: int) -> dict:
        H1, H2 = self.hidden_dims
        off = 0
        lay = {"N": N, "H1": H1, "H2": H2}
        if self.use_bias:
            lay["Dense_0/bias"] = (off, (H1,))
            off += H1
        lay["Dense_0/kernel"] = (off, (N, H1))
        off += N * H1
        if self.use_bias:
            lay["Dense_1/bias"] = (off, (H2,))
            off += H2
        lay["Dense_1/kernel"] = (off, (H1, H2))
        off += H1 * H2
        lay["n_params"] = off
        return lay

    _NAMES = ("Dense_0/bias", "Dense_0/kernel", "Dense_1/bias", "Dense_1/kernel")

    def torch_dtype(self) -> torch.dtype:
        return _ops.torch_dtype(self.param_dtype)

    def init(self, seed: int, N: int) -> torch.Tensor:
        """Flat buffer with entries N(0, std^2), one numpy Philox stream per block (ravel order)."""
        lay = self.layout(N)
        flat = np.empty(lay["n_params"], dtype=self.param_dtype)
        for k, name in enumerate(self._NAMES):
            if name in lay:
                off, shape = lay[name]
                n = int(np.prod(shape))
                std = self.bias_init_std if name.endswith("bias") else self.kernel_init_std
                rng = np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), 16 + k]))
                flat[off : off + n] = rng.normal(0.0, std, size=n).astype(self.param_dtype)
        return torch.as_tensor(flat, device=_ops.device())

    def views(self, flat: torch.Tensor, N: int) -> dict:
        lay = self.layout(N)
        out = {}
        for name in self._NAMES:
            if name in lay:
                o, s = lay[name]
                layer, leaf = name.split("/")
                out.setdefault(layer, {})[leaf] = flat[o : o + int(np.prod(s))].view(s)
        return out

    def flatten(self, params: dict, N: int) -> torch.Tensor:
        lay = self.layout(N)
        flat = torch.empty(lay["n_params"], dtype=self.torch_dtype(), device=_ops.device())
        for name in self._NAMES:
            if name not in lay:
                continue
            layer, leaf = name.split("/")
            v = params[layer][leaf]
            o, s = lay[name]
            flat[o : o + int(np.prod(s))] = torch.as_tensor(
                np.asarray(v) if not isinstance(v, torch.Tensor) else v,
                device=flat.device).reshape(-1).to(flat.dtype)
        return flat

    def native(self, flat: torch.Tensor, N: int) -> nat.NkaFfnn:
        lay = self.layout(N)
        if flat.dtype != self.torch_dtype() or not flat.is_contiguous():
            raise ValueError("parameter buffer has the wrong dtype or is not contiguous")
        m = nat.NkaFfnn()
        m.n_visible = N
        m.n_hidden1, m.n_hidden2 = self.hidden_dims
        m.dtype = _ops.dtype_code(flat.dtype)
        m.params = flat.data_ptr()
        m.b1_off = lay["Dense_0/bias"][0] if "Dense_0/bias" in lay else -1
        m.w1_off = lay["Dense_0/kernel"][0]
        m.b2_off = lay["Dense_1/bias"][0] if "Dense_1/bias" in lay else -1
        m.w2_off = lay["Dense_1/kernel"][0]
        return m

    def apply(self, flat: torch.Tensor, sigma) -> torch.Tensor:
        N = sigma.shape[-1]
        xs = _ops.as_states(sigma, N)
        batch = xs.shape[:-1]
        xf = xs.reshape(-1, N)
        out = torch.empty(xf.shape[0], dtype=flat.dtype, device=xf.device)
        m = self.native(flat, N)
        nat.call("nka_ffnn_log_psi", _ops.C.byref(m), _ops.ptr(xf), xf.shape[0], _ops.ptr(out),
                 _ops.stream())
        return out.reshape(batch)

    def __repr__(self):
        return f"FFNN(hidden_dims={self.hidden_dims}, param_dtype={self.param_dtype})"
