"""Pins for the FP64 references the precision study measures against.

oracle.fp64_chain (branch sums + numpy's FFT + the output correction) must
equal oracle.fp64_channelize, the direct definition
sum_n h[n] x[mD-n] e^{-j2pi k (mD-n)/M} (pfb_branch.comp:9-19), and the
e4m3 decoder must give the format's defined values.
"""
import numpy as np

from helpers import rng
from oracle import oracle


def test_fp64_chain_equals_direct_definition():
    M, D, L = 64, 32, 4
    g = rng(0xF64C4A1300000001)
    n = 40 * D
    codes = g.integers(-32768, 32768, size=2 * n, dtype=np.int16)
    taps = g.standard_normal(M * L).astype(np.float32)
    first = (M * L - 1 + D - 1) // D
    blocks = (n - 1) // D - first + 1
    fast = oracle.fp64_chain(codes.view(np.uint32), taps, M, D, first, blocks)
    x = codes.astype(np.float64) * 2.0 ** -15
    direct = oracle.fp64_channelize(x[0::2] + 1j * x[1::2], taps.astype(np.float64), M, D, first,
                                    blocks)
    assert np.max(np.abs(fast - direct)) <= 1e-12 * np.max(np.abs(direct))


def test_e4m3_decoder_known_values():
    table = {0x00: 0.0, 0x01: 2.0 ** -9, 0x07: 7 * 2.0 ** -9, 0x08: 2.0 ** -6, 0x38: 1.0,
             0x3C: 1.5, 0x7E: 448.0, 0xB8: -1.0, 0x40: 2.0}
    for b, v in table.items():
        assert oracle.e4m3_decode(np.array([b]))[0] == v, hex(b)
    assert np.isnan(oracle.e4m3_decode(np.array([0x7F]))[0])
