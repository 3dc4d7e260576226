"""ctypes binding of the C ABI declared in ``include/netket_amd.h``.

The library ``netket_amd/_lib/libnetket_amd.so`` is built by
``netket_amd.build``.  Loading fails loudly: there is no CPU fallback for any
kernel of the VMC step.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libnetket_amd.so"

F32 = 0
F64 = 1
OP_ISING = 0
OP_LOCAL = 1
RULE_LOCAL = 0
RULE_EXCHANGE = 1
LAYOUT_PS = 0
LAYOUT_SP = 1
OUT_BF16X2 = 0
OUT_F32 = 1
OUT_F64 = 2


class NkaLocalOp(C.Structure):
    _fields_ = [
        ("n_ops", C.c_int),
        ("max_conn_size", C.c_int),
        ("nonzero_diagonal", C.c_int),
        ("mel_dtype", C.c_int),
        ("constant", C.c_double),
        ("mel_cutoff", C.c_double),
        ("op_sites", C.c_void_p),
        ("op_nsites", C.c_void_p),
        ("op_row_off", C.c_void_p),
        ("row_diag", C.c_void_p),
        ("row_nconn", C.c_void_p),
        ("row_conn_off", C.c_void_p),
        ("conn_mel", C.c_void_p),
        ("conn_newidx", C.c_void_p),
        ("conn_term", C.c_void_p),
    ]


class NkaRbm(C.Structure):
    _fields_ = [
        ("n_visible", C.c_int),
        ("n_hidden", C.c_int),
        ("dtype", C.c_int),
        ("reserved", C.c_int),
        ("params", C.c_void_p),
        ("bias_off", C.c_int64),
        ("kernel_off", C.c_int64),
        ("visible_off", C.c_int64),
    ]


class NkaFfnn(C.Structure):
    _fields_ = [
        ("n_visible", C.c_int),
        ("n_hidden1", C.c_int),
        ("n_hidden2", C.c_int),
        ("dtype", C.c_int),
        ("params", C.c_void_p),
        ("b1_off", C.c_int64),
        ("w1_off", C.c_int64),
        ("b2_off", C.c_int64),
        ("w2_off", C.c_int64),
    ]


class NkaSweepJob(C.Structure):
    _fields_ = [
        ("model", NkaRbm),
        ("sigma", C.c_void_p),
        ("log_prob", C.c_void_p),
        ("samples", C.c_void_p),
        ("log_prob_samples", C.c_void_p),
        ("n_accepted", C.c_void_p),
        ("seed", C.c_uint64),
        ("counter", C.c_uint64),
        ("counter_dev", C.c_void_p),
    ]


class NkaIsing(C.Structure):
    _fields_ = [
        ("edges", C.c_void_p),
        ("n_edges", C.c_int),
        ("mel_dtype", C.c_int),
        ("h", C.c_double),
        ("J", C.c_double),
    ]


class NkaElocJob(C.Structure):
    _fields_ = [
        ("model", NkaRbm),
        ("op", NkaIsing),
        ("sigma", C.c_void_p),
        ("eloc", C.c_void_p),
        ("log_psi", C.c_void_p),
    ]


vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
u64 = C.c_uint64
f64 = C.c_double

# name -> (restype, argtypes); must match include/netket_amd.h
SIGNATURES = {
    "nka_version": (C.c_char_p, []),
    "nka_last_error": (C.c_char_p, []),
    "nka_ising_conn": (i32, [vp, i64, i32, vp, i32, f64, f64, i32, i32, vp, vp, vp]),
    "nka_localop_conn": (i32, [vp, i64, i32, C.POINTER(NkaLocalOp), vp, vp, vp, vp]),
    "nka_rbm_log_psi": (i32, [C.POINTER(NkaRbm), vp, i64, vp, vp]),
    "nka_ffnn_log_psi": (i32, [C.POINTER(NkaFfnn), vp, i64, vp, vp]),
    "nka_ffnn_local_energy_workspace": (i64, [C.POINTER(NkaFfnn), i32, vp, i64]),
    "nka_ffnn_local_energy": (i32, [C.POINTER(NkaFfnn), i32, vp, vp, i64, vp, vp, vp, i64, vp]),
    "nka_ffnn_metropolis": (i32, [C.POINTER(NkaFfnn), i32, vp, i32, vp, vp, vp, i32, i32, i32, i32, f64,
                                  C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp]),
    "nka_ffnn_prepare_workspace": (i64, [C.POINTER(NkaFfnn), i64]),
    "nka_ffnn_prepare": (i32, [C.POINTER(NkaFfnn), vp, i64, vp, vp, vp, vp, vp]),
    "nka_ffnn_contract_workspace": (i64, [C.POINTER(NkaFfnn), i64, i32]),
    "nka_ffnn_contract": (i32, [C.POINTER(NkaFfnn), vp, vp, vp, vp, i64, vp, i32, vp, vp, vp, vp]),
    "nka_ffnn_ntk_workspace": (i64, [C.POINTER(NkaFfnn), i64]),
    "nka_ffnn_ntk": (i32, [C.POINTER(NkaFfnn), vp, vp, vp, vp, i64, vp, f64, f64, f64, vp, i64, vp, vp]),
    "nka_ffnn_jacobian": (i32, [C.POINTER(NkaFfnn), vp, vp, vp, vp, i64, vp, f64, i32, i32, vp, vp, i64,
                                vp]),
    "nka_rbm_local_energy_workspace": (i64, [C.POINTER(NkaRbm), i32, vp, i64]),
    "nka_rbm_local_energy": (i32, [C.POINTER(NkaRbm), i32, vp, vp, i64, vp, vp, vp, i64, vp]),
    "nka_random_states": (i32, [vp, i32, i32, i32, u64, u64, vp]),
    "nka_rbm_metropolis_batched": (i32, [vp, i32, C.POINTER(NkaRbm), i32, vp, i32, i32, i32, i32, i32, f64, vp]),
    "nka_rbm_local_energy_batchable": (i32, [C.POINTER(NkaRbm), C.POINTER(NkaIsing)]),
    "nka_rbm_local_energy_batched": (i32, [vp, i32, C.POINTER(NkaRbm), C.POINTER(NkaIsing), i64, vp]),
    "nka_rbm_metropolis": (
        i32,
        [C.POINTER(NkaRbm), i32, vp, i32, vp, vp, vp, i32, i32, i32, i32, f64, u64, u64, vp, vp, vp,
         vp, vp],
    ),
    "nka_stats_moments": (i32, [vp, i32, i32, vp, vp]),
    "nka_rbm_prepare": (i32, [C.POINTER(NkaRbm), vp, i64, vp, vp, vp]),
    "nka_rbm_contract_workspace": (i64, [C.POINTER(NkaRbm), i64, i32]),
    "nka_rbm_contract": (i32, [C.POINTER(NkaRbm), vp, vp, i64, vp, i32, vp, vp, vp, vp]),
    "nka_rbm_jacobian": (i32, [C.POINTER(NkaRbm), vp, vp, vp, i64, vp, f64, i32, i32, vp, vp, i64, vp]),
    "nka_gram_bf16x3": (i32, [vp, vp, i64, i64, i64, f64, f64, vp, i64, vp]),
    "nka_rbm_qgt_workspace": (i64, [C.POINTER(NkaRbm), i64]),
    "nka_rbm_qgt": (i32, [C.POINTER(NkaRbm), vp, vp, i64, f64, f64, vp, i64, vp, vp]),
    "nka_rbm_ntk_workspace": (i64, [C.POINTER(NkaRbm), i64]),
    "nka_rbm_ntk": (i32, [C.POINTER(NkaRbm), vp, vp, i64, vp, f64, f64, f64, vp, i64, vp, vp]),
    "nka_syrk_f64": (i32, [vp, i64, i64, i64, f64, f64, f64, vp, i64, vp]),
    "nka_gemm_nt_f64": (i32, [vp, i64, vp, i64, vp, i64, i64, i64, i64, f64, f64, vp]),
    "nka_gemv_f64": (i32, [vp, i64, i64, i64, i32, vp, f64, f64, vp, vp, vp]),
    "nka_potrf_workspace": (i64, [i64]),
    "nka_potrf_f64": (i32, [vp, i64, i64, vp, vp, vp]),
    "nka_potrs_f64": (i32, [vp, i64, i64, vp, vp, vp, vp]),
    "nka_symmetrize_f64": (i32, [vp, i64, i64, vp]),
    "nka_check_finite": (i32, [vp, i64, vp, vp]),
    "nka_cholesky_solve_noalloc_max_n": (i64, []),
    "nka_cholesky_solve_f64_noalloc": (i32, [vp, i64, i64, vp, vp, vp]),
    "nka_select_on_flag_f64": (i32, [vp, vp, vp, i64, vp]),
    "nka_sgd_update": (i32, [vp, i32, vp, i64, f64, vp]),
    "nka_centre_scale": (i32, [vp, i64, vp, f64, vp, vp]),
    "nka_rbm_jvp_workspace": (i64, [C.POINTER(NkaRbm), i64]),
    "nka_rbm_jvp": (i32, [C.POINTER(NkaRbm), vp, vp, i64, vp, vp, f64, f64, vp, vp, vp]),
    "nka_ffnn_jvp_workspace": (i64, [C.POINTER(NkaFfnn), i64]),
    "nka_ffnn_jvp": (i32, [C.POINTER(NkaFfnn), vp, vp, vp, vp, i64, vp, vp, f64, f64, vp, vp, vp]),
    "nka_axpy_f64": (i32, [i64, f64, vp, vp, vp]),
    "nka_cg_workspace": (i64, [i64]),
    "nka_cg_start": (i32, [vp, i64, f64, f64, i64, vp, vp, vp]),
    "nka_cg_start_x0": (i32, [vp, i64, vp, i64, f64, f64, i64, vp, vp, vp, vp]),
    "nka_cg_iterate": (i32, [vp, i64, i64, vp, vp, i32, vp]),
    "nka_qgt_rescale_f64": (i32, [vp, i64, i64, f64, f64, i32, vp, vp]),
    "nka_mul_div_f64": (i32, [vp, vp, i64, i32, vp, vp]),
    "nka_cg_status": (i32, [vp, i64, vp, vp]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(
                f"{LIB_PATH} is missing: build the gfx950 library first "
                "(python -m netket_amd.build). There is no CPU fallback."
            )
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str):
    if status != 0:
        msg = lib().nka_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed ({status}): {msg}")


def call(name: str, *args):
    fn = getattr(lib(), name)
    check(fn(*args), name)
