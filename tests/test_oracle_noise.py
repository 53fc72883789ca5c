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


# ---------------------------------------------------------------------------
# The line kernel (noise_reference.cpp:262-478)
# ---------------------------------------------------------------------------

from revenant_amd.dsp import (CW_GUARD_DEPTH_DB, LINE_ALE, LINE_GUARD, LINE_NOTCH,  # noqa: E402
                              LINE_STATE_GUARD, LINE_STATE_MORE, LINE_STATE_WEIGHTS,
                              design_notch, line_check, line_notch_flags, line_notch_sections,
                              line_state_size, notch_response_db, notch_section_db, set_notch,
                              spectral_check, spectral_frame_for, spectral_params,
                              spectral_state_size, spectral_tables)
from tests.noise_audio_cases import RATE, line_params, voice_audio  # noqa: E402


def _line(config, params, audio, state=None):
    out = np.ascontiguousarray(audio, dtype=np.float32).copy()
    st = np.zeros(line_state_size(config), np.float32) if state is None else state
    params.count = out.size
    line_check(config, params)
    oracle.noise_line(config, params, out, st)
    return out, st


def test_notch_flags_carry_the_section_count():
    assert line_notch_flags(0) == 0 and line_notch_flags(1) == LINE_NOTCH
    for n in range(1, 5):
        assert line_notch_sections(line_notch_flags(n)) == n
    assert line_notch_sections(LINE_ALE) == 0
    assert line_notch_sections(3 << 8) == 0


def test_line_validation_refuses_what_the_reference_refuses():
    c = _lib.LineConfig()
    good = line_params(LINE_NOTCH | LINE_ALE)
    line_check(c, good)
    for field, value in (("stride", 0), ("stride", 9), ("delay", 0), ("eps", 0.0),
                         ("flags", 1 << 5), ("delay", 2048 - 127 * 7)):
        bad = _lib.LineParams.from_buffer_copy(bytes(good))
        setattr(bad, field, value)
        with pytest.raises(_lib.RvxError):
            line_check(c, bad)
    edge = _lib.LineParams.from_buffer_copy(bytes(good))
    edge.delay = 2047 - 127 * 7
    line_check(c, edge)
    for cfg in (_lib.LineConfig(lanes=0), _lib.LineConfig(lanes=48), _lib.LineConfig(history=3000),
                _lib.LineConfig(lanes=128, taps=128), _lib.LineConfig(taps=128, history=16384)):
        with pytest.raises(_lib.RvxError):
            line_check(cfg, good)


def _gain_db_float_section(design, hz, rate=RATE):
    b0, b1, b2, a1, a2 = (float(np.float32(v)) for v in design)
    z1 = np.exp(-2j * np.pi * hz / rate)
    h = (b0 + b1 * z1 + b2 * z1 * z1) / (1 + a1 * z1 + a2 * z1 * z1)
    return 20 * np.log10(abs(h))


@pytest.mark.parametrize("hz,depth,width", [(1000.0, 40.0, 100.0), (350.0, 20.0, 50.0),
                                            (2400.0, 60.0, 200.0), (15000.0, 30.0, 400.0)])
def test_notch_design_cuts_to_its_depth_and_width(hz, depth, width):
    d = design_notch(hz, depth, width, RATE)
    assert abs(_gain_db_float_section(d, hz) + depth) < 1.0
    for edge in (hz - width / 2, hz + width / 2):
        assert -4.5 < _gain_db_float_section(d, edge) < -1.5
    far = hz + 10 * width if hz < 12000 else hz - 10 * width
    assert abs(_gain_db_float_section(d, far)) < 0.2
    assert abs(notch_section_db(d, hz, RATE) - _gain_db_float_section(d, hz)) < 1e-9
    with pytest.raises(_lib.RvxError):
        design_notch(30000.0, 40.0, 100.0, RATE)


def _measured_db(params, hz, count=96000):
    n = np.arange(count)
    tone = 0.5 * np.sin(2 * np.pi * hz * n / RATE)
    out, _ = _line(_lib.LineConfig(), params, tone)
    half = slice(count // 2, None)
    return 10 * np.log10((out[half].astype(np.float64) ** 2).sum() / (tone[half] ** 2).sum())


@pytest.mark.parametrize("hz,depth,width", [(1000.0, 40.0, 100.0), (2400.0, 20.0, 200.0),
                                            (600.0, 60.0, 50.0)])
def test_a_tone_through_the_twin_loses_what_the_response_says(hz, depth, width):
    d = design_notch(hz, depth, width, RATE)
    p = _lib.LineParams(flags=LINE_NOTCH)
    set_notch(p, 0, d)
    centre = notch_response_db(hz, depth, width, RATE, hz)
    assert abs(centre + depth) < 1.0
    assert abs(_measured_db(p, hz) - centre) < 0.5
    for edge in (hz - width / 2, hz + width / 2):
        analytic = notch_response_db(hz, depth, width, RATE, edge)
        assert -4.5 < analytic < -1.5
        assert abs(_measured_db(p, edge) - analytic) < 0.05


def test_cascade_is_its_sections_run_in_order():
    audio = voice_audio(5000, 11, RATE)
    p = line_params(line_notch_flags(4))
    both, st = _line(_lib.LineConfig(), p, audio)
    one = audio
    for s in range(4):
        q = _lib.LineParams(flags=LINE_NOTCH)
        src = p if s == 0 else p.more[s - 1]
        q.b0, q.b1, q.b2, q.a1, q.a2 = src.b0, src.b1, src.b2, src.a1, src.a2
        one, st1 = _line(_lib.LineConfig(), q, one)
        got = st[:4] if s == 0 else st[LINE_STATE_MORE + 4 * (s - 1):LINE_STATE_MORE + 4 * s]
        assert np.array_equal(got, st1[:4]), s
    assert np.array_equal(both, one)


def _kept_db(out, component):
    half = slice(out.size // 2, None)
    c = component[half].astype(np.float64)
    return 20 * np.log10(abs((out[half].astype(np.float64) * c).sum() / (c * c).sum()))


def test_the_predictor_removes_a_steady_tone_and_keeps_a_moving_one():
    count = 96000
    t = np.arange(count) / RATE
    steady = (0.3 * np.sin(2 * np.pi * 1300.0 * t)).astype(np.float32)
    f = 600.0 + 200.0 * np.sin(2 * np.pi * 4.0 * t)
    moving = (0.3 * np.sin(np.cumsum(2 * np.pi * f / RATE))).astype(np.float32)
    p = _lib.LineParams(flags=LINE_ALE, stride=7, delay=240, mu=1.0e-4,
                        leak=float(np.float32(1.0 - 3.0e-6)), eps=1.0e-10)
    out, st = _line(_lib.LineConfig(), p, steady + moving)
    assert _kept_db(out, steady) < -10.0
    assert _kept_db(out, moving) > -1.0
    assert np.abs(st[LINE_STATE_WEIGHTS:LINE_STATE_WEIGHTS + 128]).sum() > 0


def test_on_cw_the_guard_keeps_the_pitch_and_takes_a_heterodyne():
    count = 96000
    t = np.arange(count) / RATE
    wanted = (0.3 * np.sin(2 * np.pi * 700.0 * t)).astype(np.float32)
    whistle = (0.3 * np.sin(2 * np.pi * 1300.0 * t)).astype(np.float32)
    # the plan a CW receiver at pitch 700, passband +-250, gets: stride
    # floor(48000 / (2.5 * 950)) = 20 clamped to 8, delay 240, guard 2 * max(100, 50) wide
    p = _lib.LineParams(flags=LINE_ALE | LINE_GUARD, stride=8, delay=240, mu=1.0e-4,
                        leak=float(np.float32(1.0 - 3.0e-6)), eps=1.0e-10)
    set_notch(p, 0, design_notch(700.0, CW_GUARD_DEPTH_DB, 200.0, RATE), guard=True)
    guarded, st = _line(_lib.LineConfig(), p, wanted + whistle)
    assert _kept_db(guarded, wanted) > -1.0
    assert _kept_db(guarded, whistle) < -6.0
    assert st[LINE_STATE_GUARD + 2] != 0.0
    bare = _lib.LineParams.from_buffer_copy(bytes(p))
    bare.flags = LINE_ALE
    unguarded, _ = _line(_lib.LineConfig(), bare, wanted + whistle)
    assert _kept_db(unguarded, wanted) < -6.0


def test_lane_split_only_reorders_the_sum():
    """lanes = 1 (one serial sum), 32 and 128 (one tap per lane, beyond this
    kernel's one wavefront, so the twin alone) agree to float rounding: the
    split changes the order of the additions and nothing else."""
    audio = voice_audio(4000, 12, RATE)
    outs = []
    for lanes in (1, 32, 128):
        c = _lib.LineConfig(lanes=lanes)
        p = line_params(LINE_ALE)
        p.count = audio.size
        out = audio.copy()
        oracle.noise_line(c, p, out, np.zeros(24 + 128 + 2048, np.float32))
        outs.append(out)
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() < 1e-4


def test_dispatch_boundaries_do_not_change_the_stream():
    audio = voice_audio(4000, 13, RATE)
    c = _lib.LineConfig()
    whole, st_whole = _line(c, line_params(LINE_NOTCH | LINE_ALE | LINE_GUARD), audio)
    st = np.zeros(line_state_size(c), np.float32)
    pieces, at = [], 0
    for n in (100, 1, 0, 256, 700, 33, 1024, 255, 2, 1629):
        piece, st = _line(c, line_params(LINE_NOTCH | LINE_ALE | LINE_GUARD), audio[at:at + n], st)
        pieces.append(piece)
        at += n
    assert at == audio.size
    assert at == audio.size
    assert np.array_equal(np.concatenate(pieces), whole)
    assert np.array_equal(st, st_whole)


# ---------------------------------------------------------------------------
# Spectral noise reduction (noise_reference.cpp:492-735)
# ---------------------------------------------------------------------------

def test_spectral_figures_and_tables():
    assert spectral_frame_for(48000) == 512 and spectral_frame_for(8000) == 64
    assert spectral_frame_for(24000) == 256 and spectral_frame_for(10**7) == 512
    assert spectral_state_size(512) == 2 + 512 + 256 + 256 + 2 * 257
    p = spectral_params(0.5)
    assert p.alpha == 1.5 and abs(p.floor_gain - 10 ** (-12 / 20)) < 1e-7
    assert p.rise == np.float32((1 - 0.998) / (1 - 0.96))
    for n in (64, 512):
        t = spectral_tables(n)
        k = np.arange(n)
        assert np.abs(t[:n] - np.cos(2 * np.pi * k / n)).max() < 1e-7
        assert np.abs(t[n:2 * n] - np.sin(2 * np.pi * k / n)).max() < 1e-7
        assert t[n // 4] == 0.0 and t[n + n // 4] == 1.0 and t[n // 2] == -1.0
        assert t[n + n // 2] == 0.0 and t[3 * n // 4] == 0.0 and t[n + 3 * n // 4] == -1.0
        w = t[2 * n:3 * n].astype(np.float64)
        assert np.abs(w[:n // 2] ** 2 + w[n // 2:] ** 2 - 1).max() < 1e-6  # overlap-add identity
        assert np.abs(t[3 * n:] - t[2 * n:3 * n] / n).max() < 1e-9
    spectral_check(512, p)
    for frame in (32, 1024, 96):
        with pytest.raises(_lib.RvxError):
            spectral_check(frame, p)
    bad = _lib.SpectralParams.from_buffer_copy(bytes(p))
    bad.eps = 0.0
    with pytest.raises(_lib.RvxError):
        spectral_check(512, bad)


def _spectral(frame, params, audio, state=None):
    out = np.ascontiguousarray(audio, dtype=np.float32).copy()
    st = np.zeros(spectral_state_size(frame), np.float32) if state is None else state
    params.count = out.size
    oracle.noise_spectral(frame, params, spectral_tables(frame), out, st)
    return out, st


def test_unit_gain_hands_back_the_input_one_frame_late():
    n = spectral_frame_for(48000)
    p = spectral_params(0.5)
    p.alpha, p.floor_gain = 0.0, 1.0
    audio = voice_audio(10000, 0x4E4F495345000007, 48000.0)
    out, _ = _spectral(n, p, audio)
    assert np.abs(out[n:].astype(np.float64) - audio[:-n]).max() < 2e-5
    assert (out[:n // 2] == 0).all()


def test_the_floor_lands_on_the_true_noise_floor():
    n = spectral_frame_for(48000)
    p = spectral_params(0.5)
    sigma = 0.05
    audio = np.random.default_rng(0x4E4F495345000008).normal(0, sigma, 48000 * 4)
    _, st = _spectral(n, p, audio.astype(np.float32))
    bins = n // 2 + 1
    minimum = st[2 + n + n // 2 + n // 2 + bins:][:bins].astype(np.float64) * p.bias
    mean = minimum[1:-1].mean()
    truth = sigma * sigma * n / 2
    assert abs(10 * np.log10(mean / truth)) < 1.5


@pytest.mark.parametrize("strength,gaps,kept", [(0.0, (-7.0, -4.5), -3.0),
                                                 (0.5, (-13.0, -9.0), -4.0),
                                                 (1.0, (-19.0, -13.0), -5.5)])
def test_noise_between_syllables_goes_to_the_floor(strength, gaps, kept):
    """A tone gated on and off every 100 ms over white noise: in the gaps the
    output noise sits near the floor gain, -(6 + 12 s) dB, and the tone, which
    the minimum tracker cannot follow up within a syllable, keeps more of
    itself than the noise does. A steady tone would be learned as noise (the
    minimum rises with a 2.7 s time constant), hence the gating."""
    n = 512
    p = spectral_params(strength)
    count = 48000 * 3
    t = np.arange(count) / 48000.0
    gate = (np.floor(t * 5) % 2 == 0).astype(np.float64)
    tone = 0.2 * np.sin(2 * np.pi * 6000.0 * t) * gate
    noise = np.random.default_rng(21).normal(0, 0.01, count)
    out, _ = _spectral(n, p, (tone + noise).astype(np.float32))
    late = slice(count // 2, None)
    shifted = out[n:][late].astype(np.float64)
    ref = tone[:-n][late]
    kept_db = 20 * np.log10((shifted * ref).sum() / (ref * ref).sum())
    off = gate[:-n][late] == 0
    gap_db = 10 * np.log10(shifted[off].var() / noise[:-n][late][off].var())
    assert gaps[0] < gap_db < gaps[1]
    assert kept_db > kept and kept_db > gap_db + 3.0


def test_spectral_dispatch_boundaries_do_not_change_the_stream():
    n = 256
    p = spectral_params(0.7)
    audio = voice_audio(4000, 14, 24000.0)
    whole, st_whole = _spectral(n, p, audio)
    st = np.zeros(spectral_state_size(n), np.float32)
    pieces, at = [], 0
    for k in (100, 1, 0, 256, 700, 33, 1024, 255, 2, 1629):
        piece, st = _spectral(n, spectral_params(0.7), audio[at:at + k], st)
        pieces.append(piece)
        at += k
    assert at == audio.size
    assert np.array_equal(np.concatenate(pieces), whole)
    assert np.array_equal(st, st_whole)
