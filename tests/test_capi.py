"""The C-ABI library loads and exports every entry point declared in
include/netket_amd.h, and the Python binding covers them all (no GPU)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "netket_amd.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nka_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "nka_rbm_metropolis" in names and "nka_gram_bf16x3" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from netket_amd import _native

    if not _native.LIB_PATH.exists():
        pytest.fail("libnetket_amd.so is not built (python -m netket_amd.build)")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    from netket_amd import _native

    assert sorted(_native.SIGNATURES) == _declared()
    _native.lib()  # sets every argtypes entry without error
    assert _native.lib().nka_version().startswith(b"netket_amd")


def test_argument_errors_are_reported_without_gpu():
    # bad sizes are rejected on the host before any launch
    from netket_amd import _native

    lib = _native.lib()
    st = lib.nka_ising_conn(None, 4, 0, None, 0, 1.0, 1.0, 1, 1, None, None, None)
    assert st == -1
    assert b"bad sizes" in lib.nka_last_error()
    st = lib.nka_gram_bf16x3(None, None, 10, 12, 12, 0.0, 0.0, None, 10, None)
    assert st == -1


def test_noalloc_cholesky_rejects_sizes_beyond_its_scratch():
    # the size check runs on the host, before the scratch lookup or any launch
    from netket_amd import _native

    lib = _native.lib()
    nmax = lib.nka_cholesky_solve_noalloc_max_n()
    assert nmax == 32768
    fake = ctypes.c_void_p(256)  # never dereferenced: the call returns before any launch
    st = lib.nka_cholesky_solve_f64_noalloc(fake, nmax + 1, nmax + 1, fake, fake, None)
    assert st == -1
    assert b"exceeds the maximum" in lib.nka_last_error()
