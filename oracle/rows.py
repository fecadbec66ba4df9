"""Sample-row formats of the grouped gridding path (TEST INFRASTRUCTURE ONLY).

The grouped BF16 modes first copy the coil-major k-space ``y (C, N)`` into
sample-major rows of its 2C values (re, im per coil).  Two of the copies
are lossy, and this module states what they must hold, independently of
the HIP encoders, so tests can compare the device's bytes with it:

``s24``: every value rounded to nearest even at 16 significant bits (the
top three bytes of the float32 pattern), NaN kept NaN, overflow to
infinity.

``bfp15``: groups of 8 values (4 coils, the last group zero-padded), 16
bytes a group = 8 little-endian 16-bit fields; bits 1..15 of field i hold
a two's-complement mantissa m_i in [-16384, 16383] and bit 0 of field i is
bit i of the group's exponent E (0..255); value_i = m_i * 2^(E - 140).  The
encoder takes M = max |x_i|, s = max(floor(log2 M) - 13, -140) (+1 when
M * 2^-s >= 16383.5, i.e. when M would round past 16383; s = -140 for
M = 0) and m_i = round-half-even(x_i * 2^-s); a group with a NaN or an
infinity stores E = 255 and m_i = -16384 (every field 0x8001) and decodes to
-infinity.

Neither format exists in the reference; they are pinned by what the
mathematics fixes (tests/test_oracle_rows.py): the error bound
|decode(encode(x)) - x| <= 2^(s-1) <= M 2^-14, exact round trips of values
already on the group's grid, idempotence, hand-computed fields, and the
rounding cases at the ends of the range.
"""

from __future__ import annotations

import numpy as np

BFP_BIAS = 140


def bfp15_groups(C: int) -> int:
    return (C + 3) // 4


def bfp15_encode(x: np.ndarray) -> np.ndarray:
    """float32 rows (N, 2C) -> uint16 fields (N, 8 * ceil(C / 4))."""
    x = np.asarray(x, dtype=np.float32)
    N, V = x.shape
    if V % 2:
        raise ValueError("rows hold (re, im) pairs")
    ng = bfp15_groups(V // 2)
    g = np.zeros((N, ng * 8), dtype=np.float64)
    g[:, :V] = x
    g = g.reshape(N, ng, 8)
    fin = np.isfinite(g)
    ok = fin.all(axis=-1)
    a = np.where(fin, np.abs(g), 0.0)
    M = a.max(axis=-1)
    _, e = np.frexp(M)  # M = f 2^e, f in [0.5, 1): floor(log2 M) = e - 1
    s = np.where(M > 0, np.maximum(e - 1 - 13, -BFP_BIAS), -BFP_BIAS)
    s = s + (np.ldexp(M, -s) >= 16383.5)
    gv = np.where(fin, g, 0.0)
    m = np.rint(np.ldexp(gv, -s[..., None])).astype(np.int64)  # exact in float64, half to even
    if np.any(ok & ((m > 16383).any(-1) | (m < -16384).any(-1))):
        raise AssertionError("bfp15 mantissa out of range")
    E = (s + BFP_BIAS).astype(np.int64)
    bits = (E[..., None] >> np.arange(8)) & 1
    f = ((m & 0x7FFF) << 1) | bits
    f = np.where(ok[..., None], f, 0x8001)
    return f.astype(np.uint16).reshape(N, ng * 8)


def bfp15_decode(f: np.ndarray, C: int) -> np.ndarray:
    """uint16 fields (N, 8 * ceil(C / 4)) -> float32 rows (N, 2C)."""
    f = np.asarray(f, dtype=np.uint16)
    N = f.shape[0]
    ng = bfp15_groups(C)
    g = f.reshape(N, ng, 8)
    E = ((g.astype(np.int64) & 1) << np.arange(8)).sum(axis=-1)
    m = (g.view(np.int16).astype(np.int64)) >> 1
    with np.errstate(over="ignore"):
        v = np.ldexp(m.astype(np.float64), (E - BFP_BIAS)[..., None]).astype(np.float32)
    return v.reshape(N, ng * 8)[:, : 2 * C]


def bfp15_scale(x: np.ndarray) -> np.ndarray:
    """Per-group scale exponent s (N, ceil(C / 4)) of finite rows (N, 2C)."""
    f = bfp15_encode(x)
    N = f.shape[0]
    g = f.reshape(N, -1, 8).astype(np.int64)
    return ((g & 1) << np.arange(8)).sum(axis=-1) - BFP_BIAS


def s24_encode(x: np.ndarray) -> np.ndarray:
    """float32 values -> the float32 values the 24-bit rows hold."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((b & 0x7F800000) == 0x7F800000) & ((b & 0x7FFFFF) != 0)
    r = (b + 0x7F + ((b >> 8) & 1)) >> 8
    r = np.where(nan, (b >> 8) | 0x4000, r)
    return ((r << 8) & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
