"""Pins the impulse blanker's oracle and its host side
(core/dsp/noise_reference.cpp:112-226).

- design figures and the threshold (noise_reference.h:96-114);
- validation refuses what the reference refuses;
- detect: each flag against a float64 restatement of the inequality where the
  two sides differ by more than float32 rounding can move them, and a zero
  window never flags;
- apply: the blanked ring against a numpy dilation of the flags by
  [-hang, +lead], across a ring wrap;
- behaviour, tests/reference/test_noise.cpp:270-322: one five-sample impulse
  30 dB above Gaussian noise is blanked with its hang and lead and nothing the
  noise did is.
"""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import BLANK_APPLY, BLANK_DETECT, blanker_check, blanker_threshold, design_blanker
from tests.blanker_cases import (CHANNEL_BLOCKS, OUT_RING, apply_params, config_144k,
                                 impulsive_channels, wrapping_params)

SEED = 0x4E4F495345000001


def _with_pass(c, which):
    out = _lib.BlankerConfig.from_buffer_copy(bytes(c))
    out.pass_ = which
    return out


def test_design_figures():
    c = design_blanker(144_000)
    assert (c.window, c.guard, c.hang, c.lead) == (72, 4, 2, 2)
    assert design_blanker(8_000).window == 16
    assert design_blanker(48_000).window == 24
    assert design_blanker(2_400_000).window == 256
    assert blanker_threshold(12.0) == np.float32(10 ** 1.2)
    assert blanker_threshold(0.0) == 1.0


def test_validation_refuses_what_the_reference_refuses():
    c = config_144k()
    p, _ = wrapping_params()
    blanker_check(c, p, BLANK_DETECT)
    blanker_check(c, p, BLANK_APPLY)
    with pytest.raises(_lib.RvxError):
        blanker_check(c, p, 2)
    for window in (15, 257):
        bad = _lib.BlankerConfig.from_buffer_copy(bytes(c))
        bad.window = window
        with pytest.raises(_lib.RvxError):
            blanker_check(bad, p, BLANK_DETECT)
    for field, value in (("chan_mask", 4000), ("out_mask", 2048), ("threshold", 0.0),
                         ("threshold", float("nan")), ("threshold", -1.0)):
        bad = _lib.BlankerParams.from_buffer_copy(bytes(p))
        setattr(bad, field, value)
        with pytest.raises(_lib.RvxError):
            blanker_check(c, bad, BLANK_DETECT)
    # reach = hang + guard + window + lead + count must stay below the smaller ring
    edge = _lib.BlankerParams.from_buffer_copy(bytes(p))
    edge.count = OUT_RING - (c.hang + c.guard + c.window + c.lead) - 1
    blanker_check(c, edge, BLANK_DETECT)
    edge.count += 1
    with pytest.raises(_lib.RvxError):
        blanker_check(c, edge, BLANK_DETECT)


def _detect(c, p, ch):
    flags = np.zeros(p.out_mask + 1, np.float32)
    oracle.blank_detect(_with_pass(c, BLANK_DETECT), p, ch, flags)
    return flags


def test_detect_is_the_threshold_inequality():
    c = config_144k()
    ch = impulsive_channels(SEED)
    p, first = wrapping_params()
    flags = _detect(c, p, ch)
    power = (ch.astype(np.float64) ** 2).sum(1)
    decided = 0
    for i in range(p.count):
        ci = (p.chan_first + i) & p.chan_mask
        o = (p.out_first + i) & p.out_mask
        idx = p.chan_base + ((ci - c.guard - c.window + np.arange(c.window)) & p.chan_mask)
        s = power[idx].sum()
        lhs = power[p.chan_base + ci] * c.window
        rhs = float(np.float32(p.threshold)) * s
        if abs(lhs - rhs) > 1e-4 * rhs:
            assert flags[o] == (1.0 if lhs > rhs else 0.0), i
            decided += 1
    assert decided > p.count - 5
    assert 20 < int(flags.sum()) < 400
    # outside the span the flag ring is untouched
    covered = (p.out_first + np.arange(p.count)) & p.out_mask
    untouched = np.ones(OUT_RING, bool)
    untouched[covered] = False
    assert not flags[untouched].any()


def test_zero_window_never_flags():
    c = config_144k()
    ch = np.zeros((CHANNEL_BLOCKS * 3, 2), np.float32)
    p, _ = wrapping_params()
    ch[p.chan_base + ((p.chan_first + 100) & p.chan_mask)] = (1.0, 0.0)
    assert not _detect(c, p, ch).any()


def test_apply_blanks_the_dilated_flags_across_a_wrap():
    c = config_144k()
    ch = impulsive_channels(SEED)
    p, first = wrapping_params()
    flags = _detect(c, p, ch)
    a = apply_params(p, first, c)
    blanked = np.full((OUT_RING, 2), 7.0, np.float32)
    oracle.blank_apply(_with_pass(c, BLANK_APPLY), a, ch, flags, blanked)
    for i in range(a.count):
        ci = (a.chan_first + i) & a.chan_mask
        o = (a.out_first + i) & a.out_mask
        near = flags[(o + np.arange(-c.hang, c.lead + 1)) & a.out_mask]
        want = (0.0, 0.0) if near.any() else tuple(ch[a.chan_base + ci])
        assert tuple(blanked[o]) == want, i
    covered = (a.out_first + np.arange(a.count)) & a.out_mask
    assert (blanked[np.setdiff1d(np.arange(OUT_RING), covered)] == 7.0).all()
    assert (blanked[covered] == 0).all(1).sum() > 20


def test_an_impulse_against_noise_is_blanked_and_the_noise_is_left_alone():
    ring, at = 8192, 5000
    rng = np.random.default_rng(0x4E4F495345000002)
    ch = rng.normal(0.0, 0.01, (ring, 2)).astype(np.float32)
    ch[at:at + 5] = (0.3, 0.1)
    c = config_144k()
    p = _lib.BlankerParams(0, ring - 1, 1000, ring - 1, 1000, 6000, blanker_threshold(12.0))
    flags = np.zeros(ring, np.float32)
    oracle.blank_detect(_with_pass(c, BLANK_DETECT), p, ch, flags)
    a = _lib.BlankerParams(0, ring - 1, 1000 + c.hang, ring - 1, 1000 + c.hang,
                           6000 - c.hang - c.lead, p.threshold)
    blanked = np.zeros((ring, 2), np.float32)
    oracle.blank_apply(_with_pass(c, BLANK_APPLY), a, ch, flags, blanked)
    span = np.arange(a.chan_first, a.chan_first + a.count)
    zero = (blanked[span] == 0).all(1)
    assert (blanked[at:at + 5] == 0).all()
    away = zero & ((span + c.lead < at) | (span > at + 4 + c.hang))
    assert int(zero.sum()) == 5 + c.hang + c.lead
    assert not away.any()
