"""Device plumbing between the NetKet-style host API and the C ABI.

PyTorch is used only for device memory and the current HIP stream; every
computation is a call into ``libnetket_amd.so``.  Functions here take and
return torch tensors resident on the GPU.
"""

from __future__ import annotations

import contextlib

import ctypes as C
from contextlib import contextmanager

import numpy as np
import torch

from . import _native as nat


class DeviceUnavailable(RuntimeError):
    pass


class KernelTimer:
    """Per-stage GPU time from HIP events recorded on the launch stream.

    Disabled by default; when enabled, every wrapped region records a start
    and an end event around its kernel launches (no synchronisation).
    """

    def __init__(self):
        self.enabled = False
        self._events = {}

    @contextmanager
    def region(self, name: str):
        if not self.enabled:
            yield
            return
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        try:
            yield
        finally:
            e.record()
            self._events.setdefault(name, []).append((s, e))

    def reset(self):
        self._events = {}

    def summary(self) -> dict:
        torch.cuda.synchronize()
        out = {}
        for name, evs in self._events.items():
            ts = [s.elapsed_time(e) for s, e in evs]
            out[name] = {"count": len(ts), "mean_ms": float(np.mean(ts)), "total_ms": float(np.sum(ts))}
        return out


TIMER = KernelTimer()


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise DeviceUnavailable(
            "netket_amd needs an MI355X GPU (torch.cuda is not available); "
            "the VMC kernels have no CPU fallback"
        )
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(0)
    return C.c_void_p(t.data_ptr())


def dtype_code(dt) -> int:
    dt = torch.empty(0, dtype=dt).dtype if not isinstance(dt, torch.dtype) else dt
    if dt == torch.float32:
        return nat.F32
    if dt == torch.float64:
        return nat.F64
    raise TypeError(f"unsupported dtype {dt}")


def torch_dtype(np_dtype) -> torch.dtype:
    np_dtype = np.dtype(np_dtype)
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}[np_dtype]


def as_states(x, n_sites: int) -> torch.Tensor:
    """Configurations as a contiguous int8 device tensor of shape (..., n_sites)."""
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev)
    else:
        t = torch.as_tensor(np.asarray(x), device=dev)
    if t.shape[-1] != n_sites:
        raise ValueError(f"configurations have {t.shape[-1]} sites, expected {n_sites}")
    if t.dtype != torch.int8:
        if t.is_floating_point():
            t = torch.round(t)
        t = t.to(torch.int8)
    return t.contiguous()


# ---------------------------------------------------------------------------
# operator tables on the device
# ---------------------------------------------------------------------------
def _ising_device(op):
    dev = device()
    key = ("ising", dev)
    c = op._device_cache.get(key)
    if c is None:
        edges = torch.as_tensor(np.ascontiguousarray(op.edges, dtype=np.int32), device=dev)
        st = nat.NkaIsing()
        st.edges = edges.data_ptr() if edges.numel() else 0
        st.n_edges = int(op.edges.shape[0])
        st.mel_dtype = nat.F64 if op.dtype == np.float64 else nat.F32
        st.h = float(op.h)
        st.J = float(op.J)
        c = {"edges": edges, "struct": st}
        op._device_cache[key] = c
    return c


def _localop_device(op):
    dev = device()
    key = ("local", dev)
    c = op._device_cache.get(key)
    if c is None:
        pk = op.pack()
        mel_t = torch_dtype(op.dtype)

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            return torch.as_tensor(a if a.size else np.zeros(1, np.int32), device=dev)

        def fl(a):
            a = np.ascontiguousarray(a, dtype=op.dtype)
            return torch.as_tensor(a if a.size else np.zeros(1, op.dtype), device=dev, dtype=mel_t)

        conn_term = np.zeros(len(pk["conn_mel"]), dtype=np.int32)
        for t in range(pk["n_ops"]):
            r0 = pk["op_row_off"][t]
            for r in range(2 ** int(pk["op_nsites"][t])):
                off = pk["row_conn_off"][r0 + r]
                conn_term[off : off + pk["row_nconn"][r0 + r]] = t
        bufs = {
            "op_sites": i32(pk["op_sites"]),
            "op_nsites": i32(pk["op_nsites"]),
            "op_row_off": i32(pk["op_row_off"]),
            "row_diag": fl(pk["row_diag"]),
            "row_nconn": i32(pk["row_nconn"]),
            "row_conn_off": i32(pk["row_conn_off"]),
            "conn_mel": fl(pk["conn_mel"]),
            "conn_newidx": i32(pk["conn_newidx"]),
            "conn_term": i32(conn_term),
        }
        st = nat.NkaLocalOp()
        st.n_ops = pk["n_ops"]
        st.max_conn_size = max(1, pk["max_conn_size"])
        st.nonzero_diagonal = int(pk["nonzero_diagonal"])
        st.mel_dtype = dtype_code(mel_t)
        st.constant = pk["constant"]
        st.mel_cutoff = pk["mel_cutoff"]
        for k, v in bufs.items():
            setattr(st, k, v.data_ptr())
        c = {"bufs": bufs, "struct": st, "packed": pk}
        op._device_cache[key] = c
    return c


def op_struct(op):
    """(kind, pointer to the C struct, keep-alive) for the fused E_loc kernel."""
    from .operator import Ising, LocalOperator

    if isinstance(op, Ising):
        c = _ising_device(op)
        return nat.OP_ISING, C.cast(C.pointer(c["struct"]), C.c_void_p), c
    if isinstance(op, LocalOperator):
        c = _localop_device(op)
        return nat.OP_LOCAL, C.cast(C.pointer(c["struct"]), C.c_void_p), c
    raise NotImplementedError(f"operator type {type(op).__name__} is not supported on the GPU path")


def ising_conn(op, x):
    N = op.hilbert.size
    xs = as_states(x, N)
    batch = xs.shape[:-1]
    flat = xs.reshape(-1, N)
    B = flat.shape[0]
    K = op._conn_rows
    mel_t = torch_dtype(op.dtype)
    xp = torch.empty((B, K, N), dtype=torch.int8, device=xs.device)
    mels = torch.empty((B, K), dtype=mel_t, device=xs.device)
    c = _ising_device(op)
    nat.call(
        "nka_ising_conn", ptr(flat), B, N, c["edges"].data_ptr() if op.edges.size else None,
        int(op.edges.shape[0]), float(op.h), float(op.J), dtype_code(mel_t), K, ptr(xp), ptr(mels),
        stream(),
    )
    return xp.reshape(*batch, K, N), mels.reshape(*batch, K)


def ising_n_conn(op, x):
    xp, mels = ising_conn(op, x)
    n = torch.full(mels.shape[:-1], op.hilbert.size if op.h != 0 else 0, dtype=torch.int32,
                   device=mels.device)
    return n + (mels[..., 0] != 0).to(torch.int32)


def localop_conn(op, x, return_n_conn: bool = False):
    N = op.hilbert.size
    xs = as_states(x, N)
    batch = xs.shape[:-1]
    flat = xs.reshape(-1, N)
    B = flat.shape[0]
    c = _localop_device(op)
    K = c["struct"].max_conn_size
    mel_t = torch_dtype(op.dtype)
    xp = torch.empty((B, K, N), dtype=torch.int8, device=xs.device)
    mels = torch.empty((B, K), dtype=mel_t, device=xs.device)
    nconn = torch.empty((B,), dtype=torch.int32, device=xs.device)
    nat.call("nka_localop_conn", ptr(flat), B, N, C.byref(c["struct"]), ptr(xp), ptr(mels),
             ptr(nconn), stream())
    xp = xp.reshape(*batch, K, N)
    mels = mels.reshape(*batch, K)
    if return_n_conn:
        return xp, mels, nconn.reshape(batch)
    return xp, mels


# ---------------------------------------------------------------------------
# linear algebra helpers
# ---------------------------------------------------------------------------
class Workspace:
    """Grow-only device scratch buffers keyed by name and stream, so that work
    enqueued on different streams never shares scratch memory."""

    def __init__(self):
        self._bufs = {}

    def get(self, name: str, nbytes: int) -> torch.Tensor:
        nbytes = max(int(nbytes), 16)
        key = (name, torch.cuda.current_stream().cuda_stream)
        b = self._bufs.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(nbytes, dtype=torch.uint8, device=device())
            self._bufs[key] = b
        return b

    def clear(self):
        self._bufs.clear()


WORKSPACE = Workspace()


@contextlib.contextmanager
def workspace_scope(ws: Workspace):
    """Route every scratch request of the enclosed calls to ``ws``.  A captured
    HIP graph keeps the raw addresses of the scratch buffers it was recorded
    with, so the capturing driver owns its buffers for the graph's lifetime
    instead of sharing the grow-only global cache, whose buffers are replaced
    (and freed) when a later request on the same stream handle needs more."""
    global WORKSPACE
    prev = WORKSPACE
    WORKSPACE = ws
    try:
        yield ws
    finally:
        WORKSPACE = prev


def potrf_(a: torch.Tensor, info: torch.Tensor, ws: torch.Tensor | None = None):
    n = a.shape[0]
    if ws is None:
        ws = WORKSPACE.get("potrf", nat.lib().nka_potrf_workspace(n))
    with TIMER.region("potrf"):
        nat.call("nka_potrf_f64", ptr(a), n, a.stride(0), ptr(ws), ptr(info), stream())
    return ws


def potrs_(a: torch.Tensor, ws: torch.Tensor, b: torch.Tensor, info: torch.Tensor):
    n = a.shape[0]
    with TIMER.region("potrs"):
        nat.call("nka_potrs_f64", ptr(a), n, a.stride(0), ptr(ws), ptr(b), ptr(info), stream())


def symmetrize_(a: torch.Tensor):
    nat.call("nka_symmetrize_f64", ptr(a), a.shape[0], a.stride(0), stream())


def gemv(a: torch.Tensor, x: torch.Tensor, trans: int = 0, alpha: float = 1.0, beta: float = 0.0,
         out: torch.Tensor | None = None) -> torch.Tensor:
    m, n = a.shape
    if out is None:
        out = torch.empty((n if trans == 1 else m,), dtype=torch.float64, device=a.device)
    ws = None
    if trans == 1:
        parts = (m + 255) // 256
        ws = WORKSPACE.get("gemv_t", parts * n * 8)
    nat.call("nka_gemv_f64", ptr(a), a.stride(0), m, n, trans, ptr(x), alpha, beta, ptr(out),
             ptr(ws), stream())
    return out


def check_finite(x: torch.Tensor, flag: torch.Tensor):
    nat.call("nka_check_finite", ptr(x), x.numel(), ptr(flag), stream())


def select_on_flag(flag: torch.Tensor, src: torch.Tensor | None, x: torch.Tensor):
    """x <- src (or 0 when src is None) if flag[0] != 0, else unchanged; on the device."""
    nat.call("nka_select_on_flag_f64", ptr(flag), ptr(src), ptr(x), x.numel(), stream())


def qgt_rescale_(a: torch.Tensor, offset: float, shift: float, mode: int, scale: torch.Tensor):
    """Scale-invariant regularisation of the lower triangle of ``a`` in place;
    ``scale`` receives sqrt(S_kk + offset)."""
    nat.call("nka_qgt_rescale_f64", ptr(a), a.shape[0], a.stride(0), float(offset), float(shift), int(mode),
             ptr(scale), stream())


def mul_div(x: torch.Tensor, d: torch.Tensor, divide: bool) -> torch.Tensor:
    """x / d or x * d as a new FP64 vector."""
    out = torch.empty_like(x)
    nat.call("nka_mul_div_f64", ptr(x), ptr(d), x.numel(), 1 if divide else 0, ptr(out), stream())
    return out
