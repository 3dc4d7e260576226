"""Device-side building blocks of the VMC step, shared by MCState, the QGT and
the drivers.  Each function enqueues C-ABI kernels on the current stream and
returns device tensors; nothing here synchronises.
"""

from __future__ import annotations

import ctypes as C
import math

import torch

from . import _native as nat
from . import _ops


def local_energies(model, flat, op, sigma2d: torch.Tensor, log_psi: bool = False):
    """E_loc (FP64) of every row of ``sigma2d`` (netket/vqs/mc/kernels.py:62-71)."""
    n, N = sigma2d.shape
    eloc = torch.empty(n, dtype=torch.float64, device=sigma2d.device)
    lp = torch.empty(n, dtype=flat.dtype, device=sigma2d.device) if log_psi else None
    kind, op_ptr, keep = _ops.op_struct(op)
    m = model.native(flat, N)
    with _ops.TIMER.region("local_energy"):
        if model.kind == "ffnn":
            wsb = nat.lib().nka_ffnn_local_energy_workspace(C.byref(m), kind, op_ptr, n)
            ws = _ops.WORKSPACE.get("ffnn_eloc", wsb)
            nat.call("nka_ffnn_local_energy", C.byref(m), kind, op_ptr, _ops.ptr(sigma2d), n,
                     _ops.ptr(eloc), _ops.ptr(lp), _ops.ptr(ws), ws.numel(), _ops.stream())
        else:
            wsb = nat.lib().nka_rbm_local_energy_workspace(C.byref(m), kind, op_ptr, n)
            ws = _ops.WORKSPACE.get("rbm_eloc", wsb) if wsb > 0 else None
            nat.call("nka_rbm_local_energy", C.byref(m), kind, op_ptr, _ops.ptr(sigma2d), n,
                     _ops.ptr(eloc), _ops.ptr(lp), _ops.ptr(ws), 0 if ws is None else ws.numel(),
                     _ops.stream())
    del keep
    return (eloc, lp) if log_psi else eloc


class FfnnFactors:
    """Per-sample factors of the network's log-derivatives (model dtype):
    h1 = log_cosh(theta1), d1 = d log psi / d theta1, g2 = tanh(z2)."""

    def __init__(self, h1, d1, g2):
        self.h1, self.d1, self.g2 = h1, d1, g2


def prepare(model, flat, sigma2d: torch.Tensor, row_major: bool = True, transposed: bool = False):
    """Per-sample factors of the log-derivatives.  RBM: tanh(theta) as (ns, M)
    and/or (M, ns) in the model dtype; FFNN: (FfnnFactors, None)."""
    n, N = sigma2d.shape
    if model.kind == "ffnn":
        H1, H2 = model.hidden_dims
        dev = sigma2d.device
        h1 = torch.empty((n, H1), dtype=flat.dtype, device=dev)
        d1 = torch.empty((n, H1), dtype=flat.dtype, device=dev)
        g2 = torch.empty((n, H2), dtype=flat.dtype, device=dev)
        m = model.native(flat, N)
        ws = _ops.WORKSPACE.get("ffnn_prepare", nat.lib().nka_ffnn_prepare_workspace(C.byref(m), n))
        with _ops.TIMER.region("prepare"):
            nat.call("nka_ffnn_prepare", C.byref(m), _ops.ptr(sigma2d), n, _ops.ptr(h1), _ops.ptr(d1),
                     _ops.ptr(g2), _ops.ptr(ws), _ops.stream())
        return FfnnFactors(h1, d1, g2), None
    M = model.n_hidden(N)
    t = torch.empty((n, M), dtype=flat.dtype, device=sigma2d.device) if row_major else None
    tT = torch.empty((M, n), dtype=flat.dtype, device=sigma2d.device) if transposed else None
    m = model.native(flat, N)
    with _ops.TIMER.region("prepare"):
        nat.call("nka_rbm_prepare", C.byref(m), _ops.ptr(sigma2d), n, _ops.ptr(t), _ops.ptr(tT),
                 _ops.stream())
    return t, tT


def contract(model, flat, sigma2d: torch.Tensor, t: torch.Tensor, w: torch.Tensor,
             mean: torch.Tensor | None = None) -> torch.Tensor:
    """sum_s (O[s,p] - mean[p]) w[v,s] for each row v of ``w`` (FP64, (n_w, Np))."""
    n, N = sigma2d.shape
    w = w.reshape(-1, n).contiguous()
    n_w = w.shape[0]
    m = model.native(flat, N)
    np_ = model.layout(N)["n_params"]
    out = torch.empty((n_w, np_), dtype=torch.float64, device=sigma2d.device)
    if model.kind == "ffnn":
        f = t
        ws = _ops.WORKSPACE.get("ffnn_contract", nat.lib().nka_ffnn_contract_workspace(C.byref(m), n, n_w))
        with _ops.TIMER.region("contract"):
            nat.call("nka_ffnn_contract", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(f.h1), _ops.ptr(f.d1),
                     _ops.ptr(f.g2), n, _ops.ptr(w), n_w, _ops.ptr(mean), _ops.ptr(out), _ops.ptr(ws),
                     _ops.stream())
        return out
    ws_bytes = nat.lib().nka_rbm_contract_workspace(C.byref(m), n, n_w)
    ws = _ops.WORKSPACE.get("contract", ws_bytes)
    with _ops.TIMER.region("contract"):
        nat.call("nka_rbm_contract", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(t), n, _ops.ptr(w),
                 n_w, _ops.ptr(mean), _ops.ptr(out), _ops.ptr(ws), _ops.stream())
    return out


def _ld(k: int) -> int:
    return ((k + 63) // 64) * 64


class JacobianBuffers:
    """Centred, rescaled Jacobian X (row-major, rows padded to a multiple of 64
    columns with zeros) in bf16 hi/lo planes or in FP64."""

    def __init__(self, rows: int, cols: int, kind: int, device):
        self.rows, self.cols, self.kind = rows, cols, kind
        self.ld = _ld(cols)
        if kind == nat.OUT_BF16X2:
            self.hi = torch.zeros((rows, self.ld), dtype=torch.uint16, device=device)
            self.lo = torch.zeros((rows, self.ld), dtype=torch.uint16, device=device)
        elif kind == nat.OUT_F64:
            self.hi = torch.zeros((rows, self.ld), dtype=torch.float64, device=device)
            self.lo = None
        else:
            self.hi = torch.zeros((rows, self.ld), dtype=torch.float32, device=device)
            self.lo = None

    def matches(self, rows, cols, kind):
        return self.rows == rows and self.cols == cols and self.kind == kind


def jacobian(model, flat, sigma2d, t, tT, mean, scale: float, layout: int, kind: int,
             bufs: JacobianBuffers | None = None) -> JacobianBuffers:
    """X = O_L^T (layout PS, Np x ns) or O_L (layout SP, ns x Np), O_L = (O - mean) scale."""
    n, N = sigma2d.shape
    np_ = model.layout(N)["n_params"]
    rows, cols = (np_, n) if layout == nat.LAYOUT_PS else (n, np_)
    if bufs is None or not bufs.matches(rows, cols, kind):
        bufs = JacobianBuffers(rows, cols, kind, sigma2d.device)
    m = model.native(flat, N)
    if model.kind == "ffnn":
        f = t
        with _ops.TIMER.region("jacobian"):
            nat.call("nka_ffnn_jacobian", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(f.h1), _ops.ptr(f.d1),
                     _ops.ptr(f.g2), n, _ops.ptr(mean), float(scale), layout, kind, _ops.ptr(bufs.hi),
                     _ops.ptr(bufs.lo), bufs.ld, _ops.stream())
        return bufs
    with _ops.TIMER.region("jacobian"):
        nat.call("nka_rbm_jacobian", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(t), _ops.ptr(tT), n,
                 _ops.ptr(mean), float(scale), layout, kind, _ops.ptr(bufs.hi),
                 _ops.ptr(bufs.lo), bufs.ld, _ops.stream())
    return bufs


def gram(bufs: JacobianBuffers, shift: float, offset: float = 0.0,
         out: torch.Tensor | None = None) -> torch.Tensor:
    """X X^T + shift I + offset 11^T on the lower triangle (FP64)."""
    n = bufs.rows
    if out is None or out.shape != (n, n):
        out = torch.empty((n, n), dtype=torch.float64, device=bufs.hi.device)
    if bufs.kind == nat.OUT_BF16X2:
        with _ops.TIMER.region("gram_bf16x3"):
            nat.call("nka_gram_bf16x3", _ops.ptr(bufs.hi), _ops.ptr(bufs.lo), n, bufs.ld, bufs.ld,
                     float(shift), float(offset), _ops.ptr(out), out.stride(0), _ops.stream())
    elif bufs.kind == nat.OUT_F64:
        if offset != 0.0:
            raise NotImplementedError("offset with FP64 Jacobians")
        with _ops.TIMER.region("gram_f64"):
            nat.call("nka_syrk_f64", _ops.ptr(bufs.hi), n, bufs.ld, bufs.ld, 1.0, 0.0,
                     float(shift), _ops.ptr(out), out.stride(0), _ops.stream())
    else:
        raise NotImplementedError("Gram of an FP32 Jacobian")
    return out


def qgt_structured_available(model, flat) -> bool:
    """Whether the dense S of ``model`` can be formed from its layer structure
    (FP32 RBM: spin-pair x activation-pair products)."""
    return model.kind == "rbm" and flat.dtype == torch.float32


def qgt_structured(model, flat, sigma2d: torch.Tensor, tT: torch.Tensor, scale: float, shift: float,
                   out: torch.Tensor | None = None) -> torch.Tensor:
    """scale (O - <O>)^T (O - <O>) + shift I on the lower triangle (FP64,
    n_params x n_params) from the RBM's Khatri-Rao structure, without a
    Jacobian; ``tT`` holds the (n_hidden, n_samples) activations and <O> the
    sample means."""
    n, N = sigma2d.shape
    np_ = model.layout(N)["n_params"]
    if out is None or out.shape != (np_, np_):
        out = torch.empty((np_, np_), dtype=torch.float64, device=sigma2d.device)
    m = model.native(flat, N)
    ws = _ops.WORKSPACE.get("rbm_qgt", nat.lib().nka_rbm_qgt_workspace(C.byref(m), n))
    with _ops.TIMER.region("qgt_structured"):
        nat.call("nka_rbm_qgt", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(tT), n, float(scale), float(shift),
                 _ops.ptr(out), out.stride(0), _ops.ptr(ws), _ops.stream())
    return out


def ntk_structured(model, flat, sigma2d: torch.Tensor, t, mean: torch.Tensor, scale: float,
                   shift: float, offset: float = 0.0, out: torch.Tensor | None = None) -> torch.Tensor:
    """scale (O - mean)(O - mean)^T + shift I + offset 11^T on the lower triangle
    (FP64) from the model's layer structure, without a Jacobian."""
    n, N = sigma2d.shape
    if out is None or out.shape != (n, n):
        out = torch.empty((n, n), dtype=torch.float64, device=sigma2d.device)
    m = model.native(flat, N)
    if model.kind == "ffnn":
        f = t
        ws = _ops.WORKSPACE.get("ffnn_ntk", nat.lib().nka_ffnn_ntk_workspace(C.byref(m), n))
        with _ops.TIMER.region("ntk_structured"):
            nat.call("nka_ffnn_ntk", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(f.h1), _ops.ptr(f.d1),
                     _ops.ptr(f.g2), n, _ops.ptr(mean), float(scale), float(shift), float(offset),
                     _ops.ptr(out), out.stride(0), _ops.ptr(ws), _ops.stream())
        return out
    ws = _ops.WORKSPACE.get("rbm_ntk", nat.lib().nka_rbm_ntk_workspace(C.byref(m), n))
    with _ops.TIMER.region("ntk_structured"):
        nat.call("nka_rbm_ntk", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(t), n, _ops.ptr(mean),
                 float(scale), float(shift), float(offset), _ops.ptr(out), out.stride(0),
                 _ops.ptr(ws), _ops.stream())
    return out


rbm_ntk = ntk_structured


def jvp(model, flat, sigma2d: torch.Tensor, t, v: torch.Tensor, mean: torch.Tensor | None,
        alpha: float, beta: float, y: torch.Tensor) -> torch.Tensor:
    """y <- alpha (O - mean) v + beta y (FP64, one entry per sample) from the
    model's layer structure, without a Jacobian."""
    n, N = sigma2d.shape
    m = model.native(flat, N)
    with _ops.TIMER.region("jvp"):
        if model.kind == "ffnn":
            f = t
            ws = _ops.WORKSPACE.get("ffnn_jvp", nat.lib().nka_ffnn_jvp_workspace(C.byref(m), n))
            nat.call("nka_ffnn_jvp", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(f.h1), _ops.ptr(f.d1),
                     _ops.ptr(f.g2), n, _ops.ptr(v), _ops.ptr(mean), float(alpha), float(beta),
                     _ops.ptr(y), _ops.ptr(ws), _ops.stream())
        else:
            ws = _ops.WORKSPACE.get("rbm_jvp", nat.lib().nka_rbm_jvp_workspace(C.byref(m), n))
            nat.call("nka_rbm_jvp", C.byref(m), _ops.ptr(sigma2d), _ops.ptr(t), n, _ops.ptr(v),
                     _ops.ptr(mean), float(alpha), float(beta), _ops.ptr(y), _ops.ptr(ws), _ops.stream())
    return y


def axpy_(y: torch.Tensor, alpha: float, x: torch.Tensor) -> torch.Tensor:
    """y += alpha x (FP64 vectors)."""
    with _ops.TIMER.region("axpy"):
        nat.call("nka_axpy_f64", y.numel(), float(alpha), _ops.ptr(x), _ops.ptr(y), _ops.stream())
    return y


def centre_scale(e: torch.Tensor, mean_dev: torch.Tensor, scale: float) -> torch.Tensor:
    out = torch.empty_like(e)
    with _ops.TIMER.region("centre"):
        nat.call("nka_centre_scale", _ops.ptr(e), e.numel(), _ops.ptr(mean_dev), float(scale),
                 _ops.ptr(out), _ops.stream())
    return out


_ZERO = {}


def scale(x: torch.Tensor, s: float) -> torch.Tensor:
    """x * s as a new FP64 vector."""
    dev = x.device
    z = _ZERO.get(dev)
    if z is None:
        z = torch.zeros(1, dtype=torch.float64, device=dev)
        _ZERO[dev] = z
    return centre_scale(x, z, s)


def sgd_update_(flat: torch.Tensor, dp: torch.Tensor, lr: float):
    with _ops.TIMER.region("sgd"):
        nat.call("nka_sgd_update", _ops.ptr(flat), _ops.dtype_code(flat.dtype), _ops.ptr(dp),
                 flat.numel(), float(lr), _ops.stream())


def inv_sqrt(n: int) -> float:
    return 1.0 / math.sqrt(n)
