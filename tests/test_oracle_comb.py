"""Pins the comb kernels' oracle and host helpers after
tests/reference/test_comb_kernels.cpp:212-437: comb_step's rounding and
wrap, the table, a pure tone projecting to amplitude times the sub-block,
subtracting a comb with its true amplitudes leaving the table's quantisation
floor, no lines copying the ring and removing -0, a subtract across the
ring's end equal to the same subtract unwrapped, validation; plus the
projection against float64 sums of the same products.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import comb_check, comb_step, comb_table
from tests.comb_lines import MASK, RATE, RING, exact_tone, make_comb, noise, wrapping_params


def test_comb_step_rounds_and_wraps():
    assert comb_step(0.0, RATE) == 0
    assert comb_step(600_000.0, RATE) == 1 << 30
    assert comb_step(-600_000.0, RATE) == 3 << 30
    assert comb_step(1_200_000.0, RATE) == 1 << 31
    assert comb_step(-1_200_000.0, RATE) == 1 << 31
    assert comb_step(0.0006, RATE) == 1
    assert comb_step(-0.0006, RATE) == 0xFFFFFFFF
    assert comb_step(1000.0, 0) == 0


def test_table_is_a_unit_circle():
    t = comb_table()
    assert t.size == 4096
    assert t[0] == np.complex64(1.0)
    assert t[1024].real == np.float32(math.cos(math.pi / 2)) and t[1024].imag == 1.0
    assert t[2048].real == -1.0
    assert np.all(np.abs(np.abs(t.astype(np.complex128)) ** 2 - 1.0) < 1e-6)


def test_pure_tone_projects_to_amplitude_times_the_sub_block():
    t = comb_table()
    line = (_lib.CombLine * 1)()
    line[0].phase0 = 0x12345678
    line[0].step = comb_step(31_415.9, RATE)
    amp = 0.4 * np.exp(1.1j)
    p = wrapping_params(1, 16)
    raw = np.zeros(RING, np.complex64)
    n = np.arange(p.count)
    raw[(p.first + n) & MASK] = (amp * exact_tone(line[0].phase0, line[0].step, n)).astype(
        np.complex64)
    proj = oracle.comb_project(p, raw, line, t)
    length = p.count // p.sub_blocks
    want = amp * length * np.exp(1j * math.pi / 4096)
    for s in range(p.sub_blocks):
        assert abs(proj[s] - want) < 2e-4 * abs(want), s


def test_subtract_with_true_amplitudes_leaves_the_quantisation_floor():
    t = comb_table()
    lines = make_comb(1024, 0xF100)
    p = wrapping_params(1024, 1)
    p.count = 16384
    n = np.arange(p.count)
    s = np.zeros(p.count, np.complex128)
    for l in range(1024):
        a = complex(lines[l].amplitude[0], lines[l].amplitude[1])
        s += a * exact_tone(lines[l].phase0, lines[l].step, n)
    power = np.sum(np.abs(s) ** 2)
    raw = np.zeros(RING, np.complex64)
    raw[(p.first + n) & MASK] = s.astype(np.complex64)

    def residual_db(with_lines):
        cleaned, _ = oracle.comb_subtract(p, raw, with_lines, t, RING)
        c = cleaned.view(np.complex64)[(p.first + n) & MASK].astype(np.complex128)
        return 10 * math.log10(np.sum(np.abs(c) ** 2) / power)

    true_db = residual_db(lines)
    lag = np.complex64(np.exp(1j * np.float32(math.pi / 4096)))
    rotated = (_lib.CombLine * 1024).from_buffer_copy(bytes(lines))
    for l in range(1024):
        a = np.complex64(complex(rotated[l].amplitude[0], rotated[l].amplitude[1])) * lag
        rotated[l].amplitude[0], rotated[l].amplitude[1] = a.real, a.imag
    tracked_db = residual_db(rotated)
    assert true_db < -60.0, true_db
    assert tracked_db < -66.0, tracked_db


def test_no_lines_copies_the_ring_and_removes_negative_zero():
    t = comb_table()
    p = wrapping_params(0, 1)
    raw = noise(RING, 7, 0.1)
    raw[(p.first + 3) & MASK] = np.complex64(complex(-0.0, -0.0))
    raw.view(np.float32)[2 * ((p.first + 3) & MASK):2 * ((p.first + 3) & MASK) + 2] = -0.0
    cleaned, removed = oracle.comb_subtract(p, raw, None, t, RING)
    slots = (p.first + np.arange(p.count)) & MASK
    rf = raw.view(np.float32).reshape(-1, 2)
    cf = cleaned.reshape(-1, 2)
    mf = removed.reshape(-1, 2)
    assert np.array_equal(cf[slots].view(np.uint32), rf[slots].view(np.uint32))
    assert np.all(mf[slots].view(np.uint32) == 0x80000000)
    assert np.array_equal((cf[slots] + mf[slots]).view(np.uint32), rf[slots].view(np.uint32))


def test_subtract_across_the_end_equals_the_same_subtract_unwrapped():
    t = comb_table()
    lines = make_comb(40, 99)
    src = noise(RING, 11, 0.05)
    w = wrapping_params(40, 1)
    f = wrapping_params(40, 1)
    f.first = 0
    n = np.arange(w.count)
    raw_w = np.zeros(RING, np.complex64)
    raw_w[(w.first + n) & MASK] = src[:w.count]
    raw_f = np.zeros(RING, np.complex64)
    raw_f[:w.count] = src[:w.count]
    cw, rw = oracle.comb_subtract(w, raw_w, lines, t, RING)
    cf, rf = oracle.comb_subtract(f, raw_f, lines, t, RING)
    slots = (w.first + n) & MASK
    assert np.array_equal(cw.view(np.uint64)[slots], cf.view(np.uint64)[n])
    assert np.array_equal(rw.view(np.uint64)[slots], rf.view(np.uint64)[n])


def test_projection_against_float64_sums():
    t = comb_table()
    lines = make_comb(8, 5)
    p = wrapping_params(8, 4)
    p.count = 4 * 1000
    raw = noise(RING, 3, 0.3)
    proj = oracle.comb_project(p, raw, lines, t)
    length = p.count // 4
    x = raw.astype(np.complex128)
    for l in range(8):
        for s in range(4):
            n = s * length + np.arange(length)
            idx = ((lines[l].phase0 + n.astype(np.uint64) * lines[l].step) % (1 << 32)) >> 20
            tone = t[idx.astype(np.int64)].astype(np.complex128)
            want = np.sum(x[(p.first + n) & MASK] * np.conj(tone))
            assert abs(proj[l * 4 + s] - want) < 1e-4 * np.sqrt(length), (l, s)


def test_validation_refuses_what_the_kernels_cannot_run():
    good = wrapping_params(8, 16)
    comb_check(good)
    for field, value in (("ring_mask", 1000), ("removed_mask", 1000), ("removed_mask", 2 * RING - 1),
                         ("lines", 1025), ("sub_blocks", 0), ("sub_blocks", 17),
                         ("count", good.count + 1), ("count", 0), ("removed_mask", 1023)):
        bad = _lib.CombParams.from_buffer_copy(bytes(good))
        setattr(bad, field, value)
        with pytest.raises(_lib.RvxError):
            comb_check(bad)
