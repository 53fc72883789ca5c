"""The bit diff can fail: the reference's proof that its harness catches a bad
kernel (tests/reference/test_divergence.cpp:26-125), on the CPU.

A harness only ever shown to pass proves nothing, so each way a result can be
wrong by the smallest amount is fed to it here and must be rejected: one unit
in the last place, a signed zero, a size mismatch, a NaN against a number. A
complex multiply with one sign flipped (the CPU twin of the suite's wrong
kernel) must be rejected against the oracle's, and only on its real parts.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_bit_exact, diff, uniform_complex


def test_one_ulp_is_caught_and_reported():
    a = np.array([1.0], dtype=np.float32)
    b = np.array([np.nextafter(np.float32(1.0), np.float32(2.0))], dtype=np.float32)
    strict = diff(b, a, seed=1234)
    assert not strict.identical
    assert strict.max_ulp_error == 1
    assert strict.first_divergence == 0
    # seed, index and both bit patterns, enough to reproduce and diagnose
    assert "0x4d2" in strict.report
    assert "0x3F800000" in strict.report and "0x3F800001" in strict.report
    # a tolerance asked for explicitly relaxes it
    assert diff(b, a, seed=1234, tolerance_ulp=1).identical
    with pytest.raises(AssertionError):
        assert_bit_exact(b, a, seed=1234)


def test_signed_zero_and_size_mismatch_are_caught():
    pos = np.array([0.0], dtype=np.float32)
    neg = np.array([-0.0], dtype=np.float32)
    assert not diff(pos, neg, seed=1).identical
    assert diff(pos, neg, seed=1).max_ulp_error == 1
    mismatch = diff(pos, np.zeros(2, dtype=np.float32), seed=1)
    assert not mismatch.identical and "size mismatch" in mismatch.report
    with pytest.raises(AssertionError, match="size mismatch"):
        assert_bit_exact(pos, np.zeros(2, dtype=np.float32))


def test_nan_matches_only_nan():
    nan_x86 = np.array([0xFFC00000], dtype=np.uint32).view(np.float32)
    nan_gfx = np.array([0x7FC00000], dtype=np.uint32).view(np.float32)
    assert diff(nan_gfx, nan_x86).identical  # vendor-chosen sign and payload
    with pytest.raises(AssertionError):
        assert_bit_exact(nan_gfx, np.array([1.0], dtype=np.float32))
    with pytest.raises(AssertionError):
        assert_bit_exact(np.array([np.inf], dtype=np.float32), nan_gfx)


def test_doubles_and_integers():
    x = np.array([1.0, -2.5], dtype=np.float64)
    y = x.copy()
    y[1] = np.nextafter(y[1], 0.0)
    res = diff(y, x)
    assert not res.identical and res.first_divergence == 1 and res.max_ulp_error == 1
    assert diff(np.array([-0.0]), np.array([0.0])).max_ulp_error == 1
    assert not diff(np.array([3, 4], np.uint32), np.array([3, 5], np.uint32)).identical
    assert diff(np.array([3, 4], np.int64), np.array([3, 4], np.int64)).identical
    assert not diff(np.array([1.0], np.float32), np.array([1.0], np.float64)).identical


def test_identical_buffers_compare_identical():
    values = uniform_complex(99, 512)
    same = diff(values, values.copy(), seed=99)
    assert same.identical and same.max_ulp_error == 0 and same.compared == 1024
    assert same.report == ""


def cmul_one_sign_flipped(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """tests/reference/shaders/cmul_wrong.comp on the CPU: the real part adds
    a.y*b.y, the imaginary part is the true product's. Each float32 operation
    rounds on its own (numpy does not contract)."""
    ar, ai = a[0::2], a[1::2]
    br, bi = b[0::2], b[1::2]
    out = np.empty_like(a)
    out[0::2] = ar * br + ai * bi
    out[1::2] = ar * bi + ai * br
    return out


def test_wrong_twin_is_rejected_on_real_parts_only():
    from oracle import oracle

    seed = 0xBADC0DE000000001
    a = uniform_complex(seed, 8192)
    b = uniform_complex(seed + 1, 8192)
    right = oracle.cmul(a, b)
    wrong = cmul_one_sign_flipped(a, b)
    res = diff(wrong, right, seed=seed)
    assert not res.identical and res.max_ulp_error > 0
    assert "seed" in res.report
    odd = diff(wrong[1::2], right[1::2])
    assert odd.identical, "the imaginary parts of the one-sign-flipped product are correct"
    assert res.first_divergence % 2 == 0
    assert not diff(wrong[0::2], right[0::2]).identical
