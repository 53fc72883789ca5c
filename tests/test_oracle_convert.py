"""Pins the oracle's format conversion against closed forms
(core/dsp/convert.h:43-105, tests/reference/test_convert.cpp:122-298)."""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib


def _run(fmt, packed, count, src_offset=0, cap=1 << 17, dst=0):
    ring = np.zeros(2 * cap, dtype=np.float32)
    oracle.convert(fmt, packed, _lib.ConvertParams(cap - 1, dst, src_offset, count), ring)
    return ring.reshape(-1, 2)


def test_cu8_every_code():
    codes = np.arange(256, dtype=np.uint8)
    out = _run(0, np.stack([codes, codes]).T.ravel().copy(), 256)[:256, 0].astype(np.float64)
    assert out[0] == -1.0 and out[255] == 1.0            # endpoints exact
    assert np.all(out == -out[::-1])                       # antisymmetric table, zero mean
    ideal = (np.arange(256) - 127.5) / 127.5
    ulp = np.spacing(np.abs(out).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(out - ideal)) <= 7.5e-8           # worst case 7.41e-08 (convert.h)
    assert np.max(np.abs(out - ideal) / ulp) <= 1.26       # 1.25 ULP at the code's magnitude
    # the scale is the float nearest 2/255 and the bias subtraction is exact
    assert np.float32(0.00784313772).view(np.uint32) == 0x3C008081
    assert np.float32(127.5).view(np.uint32) == 0x42FF0000


def test_cs8_cs16_exact_all_codes():
    c8 = np.arange(-128, 128, dtype=np.int8)
    out = _run(1, np.stack([c8, c8]).T.ravel().view(np.uint8).copy(), 256)[:256, 0]
    assert np.array_equal(out.astype(np.float64), c8 / 128.0)
    c16 = np.arange(-32768, 32768, dtype=np.int16)
    packed = np.stack([c16, c16[::-1]]).T.ravel().view(np.uint8).copy()
    out = _run(2, packed, 65536)[:65536]
    assert np.array_equal(out[:, 0].astype(np.float64), c16 / 32768.0)
    assert np.array_equal(out[:, 1].astype(np.float64), c16[::-1] / 32768.0)
    # signed zero: code 0 gives +0.0
    assert np.signbit(out[32768, 0]) == False  # noqa: E712


def test_cs24_exact_sampled():
    g = np.random.default_rng(24)
    codes = g.integers(-(1 << 23), 1 << 23, size=2 * 5001)
    codes[:4] = [-(1 << 23), (1 << 23) - 1, 0, -1]
    b = (codes & 0xFFFFFF).astype(np.uint32)
    packed = np.stack([(b >> s) & 0xFF for s in (0, 8, 16)], axis=1).astype(np.uint8).ravel()
    packed = np.concatenate([packed, np.zeros(4, np.uint8)])
    out = _run(3, packed, 5001)[:5001].ravel()
    assert np.array_equal(out.astype(np.float64), codes / float(1 << 23))


def test_offsets_and_wrap():
    """src_offset counts samples; the ring wraps modulo a power of two."""
    codes = np.arange(2 * 40, dtype=np.uint8)
    out = _run(0, codes, 10, src_offset=5, cap=16, dst=0xFFFFFFFB)
    for i in range(10):
        slot = (0xFFFFFFFB + i) & 15
        assert out[slot, 0] == np.float32((np.float32(codes[2 * (5 + i)]) - np.float32(127.5)) *
                                          np.float32(0.00784313772))


@pytest.mark.parametrize("fmt,count,nbytes", [(0, 3, 6), (2, 2, 7), (0, 20, 40)])
def test_refusals(fmt, count, nbytes):
    """A buffer short of whole words, or a count larger than the ring, is refused."""
    cap = 16
    ring = np.zeros(2 * cap, np.float32)
    with pytest.raises(ValueError):
        oracle.convert(fmt, np.zeros(nbytes, np.uint8), _lib.ConvertParams(cap - 1, 0, 0, count),
                       ring)
