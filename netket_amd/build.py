"""Build recipe for the gfx950 shared library ``netket_amd/_lib/libnetket_amd.so``.

Every ``netket_amd/csrc/*.hip`` file is compiled by ``hipcc`` for gfx950 only
and linked into one C-ABI library (declared in ``include/netket_amd.h``).
Packed-FP32 instructions are disabled for all device code.

Usage: ``python -m netket_amd.build`` (or ``build()`` from Python).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
REPO = ROOT.parent
CSRC = ROOT / "csrc"
INCLUDE = REPO / "include"
LIBDIR = ROOT / "_lib"
LIBNAME = "libnetket_amd.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-variable",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _filter(out: str) -> str:
    keep = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    return "\n".join(keep)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    hipcc = _hipcc()
    objdir = ROOT / "_build"
    objdir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")

    def compile_one(src: Path) -> Path:
        obj = objdir / (src.stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        msg = _filter(res.stdout + res.stderr)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{msg}")
        if verbose and msg.strip():
            print(msg, file=sys.stderr)
        return obj

    jobs = jobs or min(8, len(sources))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + _filter(res.stdout + res.stderr))
    out = LIBDIR / LIBNAME
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose=True)
    print(p)
