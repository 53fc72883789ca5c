"""FM stereo: the planner's resolution rule, pilot filter and rotation table, and
a stereo multiplex separated by the demodulator.

After tests/reference/test_vrx.cpp:2859-3010: a stereo station separates into
two channels by more than 30 dB, a station with no pilot comes back as two
bit-identical channels, and the decoder is off where two channels would
destroy the signal. The station is the standard pilot-tone multiplex (ITU-R
BS.450: sum, difference on a 38 kHz subcarrier in phase with a 19 kHz pilot)
written out in numpy, frequency modulated at the plan's deviation and run
through the oracle demodulator, so the pilot and rotation tables are pinned by
what they must do to a real multiplex rather than by their formulas.
"""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd.dsp import (DEEMPH_NONE, DEMOD_AM, DEMOD_CW, DEMOD_DSB, DEMOD_LSB, DEMOD_NFM,
                              DEMOD_RAW, DEMOD_USB, DEMOD_WFM, GridParams, VrxRequest, plan_vrx)

RATE = 2_400_000
GRID = GridParams(64, 16, 32)
WIDE_RATE = 122_880_000
WIDE = GridParams(512, 16, 256)   # 480 kS/s channels: a whole 200 kHz broadcast station
STATION = 20_000_000
PILOT_HZ = 19_000.0


def _wfm(**kw):
    return plan_vrx(WIDE, WIDE_RATE, VrxRequest(STATION, DEMOD_WFM, 200_000, **kw))


def _response(taps, f, rate):
    k = np.arange(taps.size)
    return np.sum(taps * np.exp(-2j * np.pi * f / rate * k))


def _pilot(p):
    d = p.demod
    base = d.audio_taps + d.dc_taps
    pairs = p.weights[base:base + 2 * d.pilot_taps].astype(np.float64)
    return pairs[0::2] + 1j * pairs[1::2]


def _rotation(p):
    d = p.demod
    base = d.audio_taps + d.dc_taps + 2 * d.pilot_taps
    pairs = p.weights[base:base + 2 * d.audio_taps].astype(np.float64)
    return pairs[0::2] + 1j * pairs[1::2]


def test_broadcast_fm_decodes_stereo_by_default():
    p = _wfm()
    assert p.plan.stereo == 1 and p.demod.channels == 2 and p.demod.pilot_taps > 16
    assert p.demod.pilot_taps % 2 == 1
    d = p.demod
    assert p.weights.size == d.audio_taps + d.dc_taps + 2 * d.pilot_taps + 2 * d.audio_taps


def test_the_pilot_filter_reaches_into_the_history():
    """The pilot filter is centred on the audio window's middle tap and reaches
    below it further than the audio filter does; the history a dispatch
    carries must cover it (test_vrx.cpp:2975-2995)."""
    p = _wfm()
    d = p.demod
    pilot_reach = d.audio_taps // 2 + d.pilot_taps - 1 + 1
    audio_reach = d.audio_taps - 1 + 1
    assert pilot_reach > audio_reach
    assert p.history >= pilot_reach


def test_composite_tap_mono_request_and_other_modes_stay_mono():
    comp = _wfm(audio_rate=171_000)
    assert comp.plan.stereo == 0 and comp.demod.channels == 1 and comp.demod.pilot_taps == 0
    mono = _wfm(stereo=False)
    assert mono.plan.stereo == 0 and mono.demod.channels == 1 and mono.demod.pilot_taps == 0
    for mode, bw in ((DEMOD_AM, 10000), (DEMOD_NFM, 12500), (DEMOD_USB, 2700),
                     (DEMOD_LSB, 2700), (DEMOD_DSB, 6000), (DEMOD_CW, 500)):
        p = plan_vrx(GRID, RATE, VrxRequest(196_500, mode, bw, cw_pitch=700))
        assert p.plan.stereo == 0 and p.demod.channels == 1 and p.demod.pilot_taps == 0
    # the raw tap is two channels for another reason: I then Q
    raw = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_RAW, 12000))
    assert raw.plan.stereo == 0 and raw.demod.channels == 2 and raw.demod.pilot_taps == 0


def test_pilot_filter_is_a_delayless_19_khz_analytic_bandpass():
    """Unit gain and zero phase at 19 kHz (the modulation cancels the tone's
    advance over the taps); the negative-frequency image, the programme band
    and the 38 kHz subcarrier are down by about the design's 70 dB."""
    p = _wfm()
    fd = p.plan.demod_rate
    h = _pilot(p)
    at_pilot = _response(h, PILOT_HZ, fd)
    assert abs(at_pilot - 1.0) < 1e-5
    assert abs(_response(h, -PILOT_HZ, fd)) < 10 ** (-65 / 20)
    for f in (0.0, 1000.0, 5000.0, 10000.0, 15000.0, 23000.0, 38000.0, 57000.0):
        assert abs(_response(h, f, fd)) < 10 ** (-65 / 20), f
    # a windowed-sinc lowpass moved to 19 kHz: symmetric about the pilot, near
    # unity close in and half amplitude at the 1 kHz cutoff
    for off in (100.0, 1000.0, 2000.0):
        up, down = abs(_response(h, PILOT_HZ + off, fd)), abs(_response(h, PILOT_HZ - off, fd))
        assert abs(up - down) < 1e-6, off
    assert abs(_response(h, PILOT_HZ + 100.0, fd)) > 0.99
    assert abs(abs(_response(h, PILOT_HZ + 1000.0, fd)) - 0.5) < 0.05


def test_rotation_carries_the_subcarrier_from_the_centre_tap():
    """Twice the unit phasor, exactly 2 at the audio window's centre, advancing
    by the 38 kHz subcarrier's angle per demodulation sample."""
    p = _wfm()
    rot = _rotation(p)
    pivot = p.demod.audio_taps // 2
    assert np.allclose(np.abs(rot), 2.0, atol=1e-6)
    assert rot[pivot] == 2.0
    step = np.angle(rot[1:] * np.conj(rot[:-1]))
    assert np.allclose(step, 2 * np.pi * 2 * PILOT_HZ / p.plan.demod_rate, atol=1e-6)


def _station(p, left_hz, right_hz, level=0.4, pilot=0.1, stereo=True, cap=1 << 16):
    """The pilot-tone multiplex in units of the plan's deviation, frequency
    modulated at the demodulation rate (a flat receiver returns it at +/-1)."""
    fd = float(p.plan.demod_rate)
    t = np.arange(cap) / fd
    left = np.sin(2 * np.pi * left_hz * t)
    right = np.sin(2 * np.pi * right_hz * t)
    wp = 2 * np.pi * PILOT_HZ * t
    m = level * (left + right) / 2
    if stereo:
        m = m + level * (left - right) / 2 * np.sin(2 * wp) + pilot * np.sin(wp)
    phase = 2 * np.pi * p.plan.deviation * np.cumsum(m) / fd
    return np.exp(1j * phase).astype(np.complex64)


def _receive(p, ring):
    cap = ring.size
    R = p.demod.decimation
    first = -(-p.history // R) + 8
    count = (cap - 1) // R - first
    dp = p.demod_block(cap - 1, first, count)
    audio = oracle.vrx_demod(p.demod, dp, ring, p.weights)
    assert audio.size == 2 * count
    t = (first + np.arange(count)) / float(p.plan.output_rate)
    return audio[0::2], audio[1::2], t


def _tones(x, t, freqs):
    cols = [np.ones_like(t)]
    for f in freqs:
        cols += [np.cos(2 * np.pi * f * t), np.sin(2 * np.pi * f * t)]
    coef, *_ = np.linalg.lstsq(np.stack(cols, 1), x.astype(np.float64), rcond=None)
    return [float(np.hypot(coef[1 + 2 * i], coef[2 + 2 * i])) for i in range(len(freqs))]


def test_a_stereo_station_separates_into_two_channels():
    p = _wfm(deemphasis=DEEMPH_NONE)
    assert p.plan.stereo == 1
    left, right, t = _receive(p, _station(p, 1000.0, 3000.0))
    l_want, l_leak = _tones(left, t, (1000.0, 3000.0))
    r_leak, r_want = _tones(right, t, (1000.0, 3000.0))
    # each channel carries its own tone at the transmitted level: a receiver
    # that handed out the sum twice would read half of it in each
    assert l_want == pytest.approx(0.4, rel=0.05)
    assert r_want == pytest.approx(0.4, rel=0.05)
    assert 20 * np.log10(l_want / l_leak) > 30.0
    assert 20 * np.log10(r_want / r_leak) > 30.0
    assert np.count_nonzero(left == right) * 20 < left.size


def test_the_curve_applies_to_both_channels():
    """The default 75 us curve on the folded filter acts on the sum and the
    difference alike, so each channel's tone comes back at the transmitted
    level times the curve (no pre-emphasis here)."""
    p = _wfm()
    left, right, t = _receive(p, _station(p, 1000.0, 3000.0))
    (l_want,) = _tones(left, t, (1000.0,))
    (r_want,) = _tones(right, t, (3000.0,))
    curve = lambda f: 1 / np.sqrt(1 + (2 * np.pi * f * 75e-6) ** 2)
    assert abs(20 * np.log10(l_want / (0.4 * curve(1000.0)))) < 0.5
    assert abs(20 * np.log10(r_want / (0.4 * curve(3000.0)))) < 0.5


def test_a_station_with_no_pilot_comes_back_as_two_identical_channels():
    p = _wfm(deemphasis=DEEMPH_NONE)
    left, right, t = _receive(p, _station(p, 1000.0, 1000.0, stereo=False))
    assert np.array_equal(left.view(np.uint32), right.view(np.uint32))
    (amp,) = _tones(left, t, (1000.0,))
    assert amp == pytest.approx(0.4, rel=0.05)
