"""Pins impulse excision's oracle and its host side against the reference's
own measurements (tests/reference/test_impulse_excision.cpp:344-477): every
impulse found and removed with a tone kept to 0.1 dB, cleaned + removed ==
raw to the bit, no sample moved; a strong clean carrier left alone; bursts
longer than the maximum width spared. Plus the design's figures
(impulse_excision.h:123-150), validation, and the median pass against numpy's
median of the same windows.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.excision import (REFERENCE, check, design_excision, excision_threshold,
                                   size_excision_rings)
from tests.excision_stream import OraclePasses, impulsive_capture, run_stream

RATE = 2_400_000


def test_design_figures():
    c = design_excision(RATE)
    assert (c.max_width, c.hang, c.lead, c.seg_log2, c.segments) == (48, 2, 2, 4, 32)
    c = design_excision(20_000_000)
    assert (c.max_width, c.hang, c.seg_log2) == (400, 2, 7)
    c = design_excision(122_880_000)
    assert (c.max_width, c.hang, c.seg_log2) == (2458, 12, 10)
    assert design_excision(10**10).max_width == 4096 and design_excision(10**10).hang == 16
    assert abs(excision_threshold(15.0) - 31.622776) < 1e-5


def test_validation_refuses_what_the_reference_refuses():
    c = design_excision(RATE)
    rings = size_excision_rings(c, 4096, 1, (1 << 15) - 1)
    good = _lib.ExcisionParams((1 << 15) - 1, 0, 16, rings.seg_mask, rings.flag_mask,
                               rings.removed_mask, 0, 31.6)
    check(c, good, REFERENCE)
    for field, value in (("ring_mask", 1000), ("first", 3), ("count", rings.seg_mask),
                         ("threshold", 0.0)):
        bad = _lib.ExcisionParams.from_buffer_copy(bytes(good))
        setattr(bad, field, value)
        with pytest.raises(_lib.RvxError):
            check(c, bad, REFERENCE)
    for field, value in (("seg_log2", 1), ("segments", 65), ("max_width", 0), ("hang", 17)):
        bc = _lib.ExcisionConfig.from_buffer_copy(bytes(c))
        setattr(bc, field, value)
        with pytest.raises(_lib.RvxError):
            check(bc, good, REFERENCE)
    with pytest.raises(_lib.RvxError):
        size_excision_rings(c, 1 << 20, 1, (1 << 15) - 1)


def test_reference_pass_is_the_median_of_the_window():
    c = design_excision(RATE)
    c.pass_ = REFERENCE
    seg = 256
    power = np.random.default_rng(5).exponential(1.0, seg).astype(np.float32)
    level = np.zeros(seg, np.float32)
    S = 1 << c.seg_log2
    first_seg = 100
    p = _lib.ExcisionParams(seg * S - 1, first_seg * S, 50, seg - 1, 0, 0, 10**6, 31.6)
    oracle.excise_reference(c, p, power, level)
    for j in range(50):
        s = first_seg + j
        window = np.sort(power[[(s - 1 - c.segments + m) % seg for m in range(c.segments)]])
        assert level[s % seg] == window[c.segments // 2]
    # before a whole window of this run: zero
    p.history = -S
    level[:] = 7.0
    oracle.excise_reference(c, p, power, level)
    assert level[first_seg % seg] == 0.0


def test_removes_impulses_keeps_the_tone_and_moves_nothing():
    length, cycles = 1 << 19, 0.0371
    where = []
    dirty = impulsive_capture(length, 20261009, 0.1, cycles, 40, 4, where)
    clean = impulsive_capture(length, 20261009, 0.1, cycles, 0, 1)
    cleaned, removed, counters, end, _ = run_stream(dirty, RATE, 65536, 15.0, OraclePasses())
    assert end > length - 1024
    assert 38 <= counters[1] <= 40 and counters[2] == 0, counters
    peak = max(np.abs(cleaned[a - 8:a + 16]).max() for a in where)
    assert peak < 0.25, peak
    n = np.arange(end)
    w = np.exp(-2j * math.pi * cycles * n)
    ref = abs(np.mean(clean[:end].astype(np.complex128) * w)) ** 2
    kept = abs(np.mean(cleaned[:end].astype(np.complex128) * w)) ** 2
    assert abs(10 * math.log10(kept / ref)) < 0.1
    total = (cleaned[:end].view(np.float32) + removed[:end].view(np.float32))
    assert np.array_equal(total.view(np.uint32), dirty[:end].view(np.float32).view(np.uint32))
    c = cleaned.view(np.float32).reshape(-1, 2)
    excised = (c[:, 0] == 0) & np.signbit(c[:, 0]) & (c[:, 1] == 0) & np.signbit(c[:, 1])
    assert excised.sum() == counters[0]
    assert np.array_equal(cleaned[~excised], dirty[~excised])


def test_leaves_a_clean_strong_carrier_alone():
    length = 1 << 18
    g = np.random.default_rng(20261010)
    n = np.arange(length)
    cap = ((g.normal(0, 0.005, length) + 1j * g.normal(0, 0.005, length)) +
           0.7 * np.exp(2j * math.pi * 0.11 * n)).astype(np.complex64)
    cleaned, _, counters, _, _ = run_stream(cap, RATE, 32768, 15.0, OraclePasses())
    assert counters[0] == 0 and counters[2] == 0
    assert np.array_equal(cleaned.view(np.uint64), cap.view(np.uint64))


def test_leaves_bursts_longer_than_the_maximum_width():
    length = 1 << 18
    c = design_excision(RATE)
    cap = impulsive_capture(length, 20261011, 0.0, 0.0, 0, 1)
    g = np.random.default_rng(20261011)
    bursts = [(100_000, 3 * c.max_width), (180_000, c.max_width + 1 + 2 * c.hang)]
    for at, n in bursts:
        cap[at:at + n] += (g.normal(0, 0.3, n) + 1j * g.normal(0, 0.3, n)).astype(np.complex64)
    cleaned, _, counters, _, _ = run_stream(cap, RATE, 65536, 15.0, OraclePasses())
    for at, n in bursts:
        lo, hi = at - 64, at + n + 64
        assert np.array_equal(cleaned[lo:hi].view(np.uint64), cap[lo:hi].view(np.uint64)), at
    assert counters[2] >= 2


def test_segment_and_flag_passes_against_numpy():
    """Segment powers are the float64 sums of |x|^2 to float rounding, and a
    sample is flagged exactly when |x|^2 S > T r (margins kept off ties)."""
    from revenant_amd.excision import FLAG, SEGMENT
    c = design_excision(RATE)
    S = 1 << c.seg_log2
    ring_n, seg_n, flag_n = 1 << 14, 1 << 10, 1 << 13
    g = np.random.default_rng(9)
    raw = (g.standard_normal(ring_n) + 1j * g.standard_normal(ring_n)).astype(np.complex64)
    raw[::97] *= 30.0
    rf = raw.view(np.float32)
    power = np.zeros(seg_n, np.float32)
    c.pass_ = SEGMENT
    first_seg, nseg = 40, 200
    p = _lib.ExcisionParams(ring_n - 1, first_seg * S, nseg, seg_n - 1, flag_n - 1, 0, 0, 31.6)
    oracle.excise_segment(c, p, rf, power)
    z = raw.astype(np.complex128)
    for j in range(nseg):
        s = first_seg + j
        want = np.sum(np.abs(z[s * S:(s + 1) * S]) ** 2)
        assert abs(power[s % seg_n] - want) <= 1e-5 * want
    level = g.exponential(2.0, seg_n).astype(np.float32)
    level[7] = 0.0  # a zero reference never flags
    flags = np.zeros(flag_n, np.uint32)
    c.pass_ = FLAG
    p = _lib.ExcisionParams(ring_n - 1, 5 * S, 3000, seg_n - 1, flag_n - 1, 0, 0,
                            excision_threshold(15.0))
    oracle.excise_flag(c, p, rf, level, flags)
    idx = 5 * S + np.arange(3000)
    pw = np.abs(z[idx]) ** 2 * S
    ref = excision_threshold(15.0) * level[(idx // S) % seg_n].astype(np.float64)
    want = (pw > ref) & (level[(idx // S) % seg_n] > 0)
    clear = np.abs(pw - ref) > 1e-4 * ref
    assert np.array_equal(flags[idx % flag_n][clear].astype(bool), want[clear])
    assert want.sum() > 10
