"""Philox4x32-10 (Salmon et al., SC'11) in NumPy — oracle for the sampler's RNG.

Test infrastructure only.  Pinned by the Random123 known-answer vectors in
tests/test_oracle_basic.py.  The draw helpers mirror the device conventions
of netket_amd/csrc/nka_common.hpp (uniform01, bounded).
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

TAG_METROPOLIS = 0x4D455452
TAG_RANDSTATE = 0x52414E44


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32; key: (2,) uint32 (broadcast). Returns (..., 4) uint32."""
    c = np.asarray(ctr, dtype=np.uint32)
    x0, x1, x2, x3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
                k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
            p0 = M0 * x0
            p1 = M1 * x2
            n0 = (p1 >> np.uint64(32)) ^ x1 ^ np.uint64(k0)
            n1 = p1 & MASK32
            n2 = (p0 >> np.uint64(32)) ^ x3 ^ np.uint64(k1)
            n3 = p0 & MASK32
            x0, x1, x2, x3 = n0, n1, n2, n3
    return np.stack([x0, x1, x2, x3], axis=-1).astype(np.uint32)


def draw(q: int, chain: int, tag: int, seed: int):
    """One 4-word draw at counter (q, chain, tag) with key = seed (64 bit)."""
    ctr = np.array([q & 0xFFFFFFFF, (q >> 32) & 0xFFFFFFFF, chain & 0xFFFFFFFF, tag], dtype=np.uint32)
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    return philox4x32_10(ctr, key)


def uniform01_f32(r) -> np.float32:
    return np.float32(int(r[1]) >> 8) * np.float32(2.0**-24)


def uniform01_f64(r) -> float:
    u = ((int(r[1]) << 32) | int(r[2])) >> 11
    return float(u) * 2.0**-53


def bounded(x: int, n: int) -> int:
    return (int(x) * int(n)) >> 32
