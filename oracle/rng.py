"""Philox-4x32-10 + Box-Muller restatement of the device init stream (test infrastructure).

The GPU engine's random_CPD replacement draws factor entries from a
Philox-4x32-10 counter stream (Salmon, Moraes, Dror, Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11): counter = (e mod 2^32, e >> 32, mode, 0)
for element e = i + I_n * r, key = (seed mod 2^32, seed >> 32), one normal per
counter by Box-Muller on two 53-bit uniforms.  Pinned by the Random123
known-answer vectors (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox-4x32-10 on uint32 counter arrays; returns four uint32 arrays."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3)]
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + _W0) & 0xFFFFFFFF
        k1 = (k1 + _W1) & 0xFFFFFFFF
    return [x.astype(np.uint32) for x in c]


def normal_matrix(I: int, R: int, seed: int, mode: int) -> np.ndarray:
    """The I x R matrix of N(0,1) draws of one mode (column-major element order)."""
    e = np.arange(I * R, dtype=np.uint64)
    x0, x1, x2, x3 = philox4x32_10(e & _MASK, e >> np.uint64(32), np.full_like(e, mode),
                                   np.zeros_like(e), seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    a = (x0.astype(np.uint64) << np.uint64(32)) | x1.astype(np.uint64)
    b = (x2.astype(np.uint64) << np.uint64(32)) | x3.astype(np.uint64)
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return z.reshape(R, I).T.copy()


def random_factors_philox(dims, R: int, seed: int):
    """random_factors (cpd.jl:48-60) on the Philox stream: draw, row_norm, lambda of last mode."""
    from .als import row_norm
    factors, lam = [], None
    for m, I in enumerate(dims):
        f, lam = row_norm(normal_matrix(int(I), int(R), int(seed), m))
        factors.append(f)
    return factors, lam
