"""Pins of the sample-row format oracle (``oracle/rows.py``), CPU only.

The block-floating-point format has no reference counterpart; these tests
fix it by hand-computed fields, the error bound the format promises, exact
round trips of on-grid values, idempotence and the ends of its range, and
cross-check the 24-bit rounding against an independent float64 statement.
"""

import numpy as np
import pytest

from oracle import rows


def _fields(E, m):
    return [((mi & 0x7FFF) << 1) | ((E >> i) & 1) for i, mi in enumerate(m)]


def test_bfp15_hand_computed_group():
    # one coil (re, im) = (1, -0.5) plus zero padding: M = 1 = 2^0, s = -13,
    # E = 127, m = (8192, -4096, 0, ...)
    x = np.array([[1.0, -0.5]], dtype=np.float32)
    f = rows.bfp15_encode(x)
    assert f.shape == (1, 8)
    assert f[0].tolist() == _fields(127, [8192, -4096, 0, 0, 0, 0, 0, 0])
    assert rows.bfp15_decode(f, 1).tolist() == [[1.0, -0.5]]


def test_bfp15_rounding_past_the_top_bumps_the_scale():
    # M * 2^13 = 16383.75 rounds past 16383: s = -12, m = round(16383.75 / 2) = 8192
    M = np.float32(16383.75 / 8192.0)
    x = np.array([[M, 0.25, -M, 0.0, 1.0, 0.0, 0.0, 0.0]], dtype=np.float32)
    s = rows.bfp15_scale(x)
    assert s.tolist() == [[-12]]
    f = rows.bfp15_encode(x)
    m = (f.view(np.int16).astype(np.int64) >> 1)[0]
    assert m.tolist() == [8192, 1024, -8192, 0, 4096, 0, 0, 0]
    # just below: 16383.25 stays at s = -13 (m = 16383)
    x[0, 0] = np.float32(16383.25 / 8192.0)
    x[0, 2] = 0.0
    assert rows.bfp15_scale(x).tolist() == [[-13]]
    assert (rows.bfp15_encode(x).view(np.int16) >> 1)[0, 0] == 16383


def test_bfp15_ties_round_to_even():
    # M = 2^13 * 2^-13 = 1 -> s = -13: 2^-14 (half a step) rounds to 0,
    # 3 * 2^-14 (1.5 steps) to 2 steps
    x = np.array([[1.0, 2.0 ** -14, 3 * 2.0 ** -14, -(2.0 ** -14), 5 * 2.0 ** -14, 0, 0, 0]], dtype=np.float32)
    m = (rows.bfp15_encode(x).view(np.int16).astype(np.int64) >> 1)[0]
    assert m.tolist() == [8192, 0, 2, 0, 2, 0, 0, 0]


@pytest.mark.parametrize("C", [1, 3, 4, 16, 32, 33])
def test_bfp15_error_bound_and_idempotence(C):
    rng = np.random.default_rng(C)
    N = 300
    x = (rng.standard_normal((N, 2 * C)) * np.exp2(rng.integers(-60, 60, (N, 1)))).astype(np.float32)
    x[::7, : min(2 * C, 5)] *= 1e-6  # groups with a large spread
    f = rows.bfp15_encode(x)
    assert f.shape == (N, 8 * ((C + 3) // 4))
    d = rows.bfp15_decode(f, C)
    s = rows.bfp15_scale(x)
    step = np.repeat(np.exp2(s.astype(np.float64)), 8, axis=1)[:, : 2 * C]
    err = np.abs(d.astype(np.float64) - x.astype(np.float64))
    assert np.all(err <= 0.5 * step)
    pad = np.zeros((N, 8 * ((C + 3) // 4)), np.float64)
    pad[:, : 2 * C] = np.abs(x)
    M = np.repeat(pad.reshape(N, -1, 8).max(-1), 8, axis=1)[:, : 2 * C]
    assert np.all(err <= M * 2.0 ** -14)
    # decoded values lie on the group's grid: encoding them again is exact
    assert np.array_equal(rows.bfp15_encode(d), f)
    assert np.array_equal(rows.bfp15_decode(rows.bfp15_encode(d), C), d)


def test_bfp15_exact_for_on_grid_values_and_range_ends():
    rng = np.random.default_rng(1)
    m = rng.integers(-16384, 16384, (50, 8))
    m[:, 0] = 16383  # the largest magnitude fixes s
    for s in (-140, -20, 0, 30, 100):
        x = np.ldexp(m.astype(np.float64), s).astype(np.float32)
        assert np.array_equal(rows.bfp15_decode(rows.bfp15_encode(x), 4), x)
        assert np.all(rows.bfp15_scale(x) == s)
    # subnormal groups use s = -140 (or -139 when they round past the top)
    sub = np.array([[1e-42, -3e-43, 0, 0, 0, 0, 0, 0]], dtype=np.float32)
    assert rows.bfp15_scale(sub).tolist() == [[-140]]
    # (the format's finest step is 2^-140: subnormal inputs round to it)
    want = (np.rint(sub.astype(np.float64) * 2.0 ** 140) * 2.0 ** -140).astype(np.float32)
    assert np.array_equal(rows.bfp15_decode(rows.bfp15_encode(sub), 4), want)
    assert want[0, :2].tolist() == [np.float32(2.0 ** -140), 0.0]
    top = np.array([[np.nextafter(np.float32(2.0 ** -126), np.float32(0)), 0, 0, 0, 0, 0, 0, 0]], np.float32)
    assert rows.bfp15_scale(top).tolist() == [[-139]]
    # the largest finite float rounds into infinity; a non-finite group decodes to -inf
    big = np.array([[np.finfo(np.float32).max, 1.0, 0, 0, 0, 0, 0, 0]], dtype=np.float32)
    assert rows.bfp15_scale(big).tolist() == [[115]]
    assert np.isposinf(rows.bfp15_decode(rows.bfp15_encode(big), 4)[0, 0])
    for bad in (np.inf, -np.inf, np.nan):
        v = np.array([[1.0, bad, 0, 0, 0, 0, 0, 0, 2.0, 3.0]], dtype=np.float32)
        f = rows.bfp15_encode(v)
        assert f[0, :8].tolist() == [0x8001] * 8
        dec = rows.bfp15_decode(f, 5)
        assert np.all(np.isneginf(dec[0, :8])) and dec[0, 8:].tolist() == [2.0, 3.0]
    zero = np.zeros((1, 8), np.float32)
    assert rows.bfp15_encode(zero)[0].tolist() == _fields(0, [0] * 8)


def test_s24_matches_a_float64_statement():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(20000) * np.exp2(rng.integers(-140, 120, 20000))).astype(np.float32)
    x[:4] = [0.0, -0.0, 1.0 + 2.0 ** -16, 1.0 + 3 * 2.0 ** -16]  # ties go to even
    got = rows.s24_encode(x)
    x64 = x.astype(np.float64)
    e = np.full(x64.shape, -126.0)
    nz = x64 != 0
    e[nz] = np.maximum(np.floor(np.log2(np.abs(x64[nz]))), -126.0)
    q = np.exp2(e - 15.0)
    with np.errstate(over="ignore"):
        want = (np.rint(x64 / q) * q).astype(np.float32)
    assert np.array_equal(got, want)
    assert got[2] == 1.0 and got[3] == 1.0 + 4 * 2.0 ** -16
