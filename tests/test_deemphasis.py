"""De-emphasis: the curve, its resolution per mode and its fold into the audio filter.

After tests/reference/test_vrx.cpp:2440-2580 (the curve resolves per mode and
never reaches a composite tap; a flat curve leaves the audio filter as it
was) and :2590-2650 (a de-emphasised receiver follows the one-pole across the
audio band, within 0.5 dB of the continuous-time prototype). The curve is
pinned by what the mathematics fixes (unit DC gain, the pole's ratio, the
shortest series that no float can tell from the infinite one, the convolution
theorem for the fold), not by restating the design.
"""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEEMPH_DEFAULT, DEEMPH_EU50, DEEMPH_NONE, DEEMPH_US75, DEMOD_AM,
                              DEMOD_CW, DEMOD_DSB, DEMOD_LSB, DEMOD_NFM, DEMOD_RAW, DEMOD_SAM,
                              DEMOD_USB, DEMOD_WFM, GridParams, VrxRequest, design_deemphasis,
                              plan_vrx)

RATE = 2_400_000
GRID = GridParams(64, 16, 32)     # 75 kS/s channels
STATION = 20_000_000
WIDE_RATE = 122_880_000
WIDE = GridParams(512, 16, 256)   # 480 kS/s channels: a whole 200 kHz broadcast station
TAU = {DEEMPH_US75: 75e-6, DEEMPH_EU50: 50e-6}
FLOOR = 2.0 ** -24
WARP_BUDGET_DB = 0.5              # test_vrx.cpp:2604
# The WFM receivers here are mono (stereo=False), as in test_vrx.cpp:2613, so
# the curve reads straight off one channel.


def _wfm(**kw):
    """A mono broadcast FM receiver on the wide grid."""
    return plan_vrx(WIDE, WIDE_RATE, VrxRequest(STATION, DEMOD_WFM, 200_000, stereo=False, **kw))


def _response(taps, f, rate):
    k = np.arange(taps.size)
    return np.sum(taps.astype(np.float64) * np.exp(-2j * np.pi * f / rate * k))


def _analog(curve, f):
    wt = 2 * np.pi * f * TAU[curve]
    return 1.0 / np.sqrt(1.0 + wt * wt)


@pytest.mark.parametrize("curve", [DEEMPH_US75, DEEMPH_EU50])
@pytest.mark.parametrize("audio_rate", [8000, 12000, 24000, 48000])
def test_curve_is_the_truncated_geometric_series(curve, audio_rate):
    taps = design_deemphasis(curve, audio_rate)
    a = np.exp(-1.0 / (audio_rate * TAU[curve]))
    n = taps.size
    # unit gain at DC, to float rounding of each tap
    assert abs(float(np.sum(taps.astype(np.float64))) - 1.0) < n * 2 ** -24
    # each tap is the last times the pole
    ratio = taps[1:].astype(np.float64) / taps[:-1].astype(np.float64)
    assert np.allclose(ratio, a, rtol=2e-6)
    # the shortest series past the float floor, plus the one tap of margin:
    # a^(n-1) is at or below 2^-24 (1 - a) and a^(n-2) is above it
    assert a ** (n - 1) <= FLOOR * (1 - a) < a ** (n - 2)
    # monotone decay, all positive
    assert np.all(taps > 0) and np.all(np.diff(taps) < 0)


def test_curve_length_is_capped_at_1024_taps():
    # 75 us at 10 MS/s wants some 13,500 terms; the series stops at 1024
    taps = design_deemphasis(DEEMPH_US75, 10_000_000)
    assert taps.size == 1024
    assert abs(float(np.sum(taps.astype(np.float64))) - 1.0) < 1e-5


def test_no_curve_is_one_unit_tap():
    assert np.array_equal(design_deemphasis(DEEMPH_NONE, 48000), np.ones(1, np.float32))
    with pytest.raises(_lib.RvxError):
        design_deemphasis(DEEMPH_DEFAULT, 48000)   # a request, not a curve
    with pytest.raises(_lib.RvxError):
        design_deemphasis(DEEMPH_US75, 0)


@pytest.mark.parametrize("curve", [DEEMPH_US75, DEEMPH_EU50])
def test_curve_follows_the_one_pole_within_the_warp_budget(curve):
    """The impulse-invariant pole is the analog one-pole aliased: within 0.5 dB
    of 1/|1 + j 2 pi f tau| over the programme band at 48 kS/s, and -3 dB
    near the corner 1/(2 pi tau)."""
    taps = design_deemphasis(curve, 48000)
    for f in (500.0, 1000.0, 2122.0, 5000.0, 8000.0):
        err = 20 * np.log10(abs(_response(taps, f, 48000)) / _analog(curve, f))
        assert abs(err) < WARP_BUDGET_DB, (curve, f, err)
    corner = 1.0 / (2 * np.pi * TAU[curve])
    assert abs(20 * np.log10(abs(_response(taps, corner, 48000))) + 3.0103) < 0.1


def test_broadcast_fm_takes_the_north_american_curve():
    p = _wfm()
    pl = p.plan
    assert pl.deemphasis == DEEMPH_US75
    n = design_deemphasis(DEEMPH_US75, pl.audio_rate).size
    assert n > 1
    # folded into the decimation filter at the audio stride: still one FIR
    assert pl.demod.audio_taps == pl.audio_decimation_taps + (n - 1) * pl.demod.decimation
    # the programme band: 15 kHz pass, stop before the 19 kHz pilot
    assert pl.audio_pass_hz == 15000.0 and pl.audio_stop_hz == 19000.0


def test_the_composite_tap_takes_no_curve():
    p = _wfm(audio_rate=171_000)
    assert p.plan.deemphasis == DEEMPH_NONE
    assert p.plan.demod.audio_taps == p.plan.audio_decimation_taps
    assert p.plan.audio_pass_hz == pytest.approx(0.4 * 171_000)


@pytest.mark.parametrize("mode,bw", [(DEMOD_AM, 10000), (DEMOD_NFM, 12500), (DEMOD_USB, 2700),
                                     (DEMOD_LSB, 2700), (DEMOD_DSB, 6000), (DEMOD_CW, 500),
                                     (DEMOD_SAM, 10000)])
def test_every_other_mode_defaults_to_none(mode, bw):
    p = plan_vrx(GRID, RATE, VrxRequest(196_500, mode, bw, cw_pitch=700))
    assert p.plan.deemphasis == DEEMPH_NONE
    assert p.plan.demod.audio_taps == p.plan.audio_decimation_taps


def test_the_raw_tap_never_carries_a_curve():
    p = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_RAW, 12000, deemphasis=DEEMPH_US75))
    assert p.plan.deemphasis == DEEMPH_NONE and p.plan.demod.audio_taps == 1


def test_an_explicit_curve_is_honoured():
    eu = _wfm(deemphasis=DEEMPH_EU50)
    assert eu.plan.deemphasis == DEEMPH_EU50
    nfm = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_NFM, 12000, deemphasis=DEEMPH_US75))
    assert nfm.plan.deemphasis == DEEMPH_US75
    n = design_deemphasis(DEEMPH_US75, 48000).size
    assert nfm.plan.demod.audio_taps == (nfm.plan.audio_decimation_taps +
                                         (n - 1) * nfm.plan.demod.decimation)
    # an explicit curve on the composite tap is honoured too
    comp = _wfm(audio_rate=171_000, deemphasis=DEEMPH_US75)
    assert comp.plan.deemphasis == DEEMPH_US75


def test_an_unknown_curve_is_refused():
    with pytest.raises(_lib.RvxError):
        _wfm(deemphasis=4)


def test_a_flat_curve_leaves_the_audio_filter_as_it_was():
    """NONE is the decimation filter itself, the table a receiver had before
    de-emphasis; the curve's own table is the same decimation filter with the curve
    folded in, whose response is the product of the two (convolution theorem)."""
    off = _wfm(deemphasis=DEEMPH_NONE)
    on = _wfm()
    dt = off.plan.audio_decimation_taps
    assert off.plan.demod.audio_taps == dt and on.plan.audio_decimation_taps == dt
    # the DC window after the audio filter is untouched by the fold
    assert np.array_equal(off.weights[dt:], on.weights[on.plan.demod.audio_taps:])
    decim = off.weights[:dt]
    folded = on.weights[:on.plan.demod.audio_taps]
    curve = design_deemphasis(DEEMPH_US75, on.plan.audio_rate)
    fd, R = on.plan.demod_rate, on.plan.demod.decimation
    for f in (0.0, 300.0, 1000.0, 2122.0, 6000.0, 14000.0, 30000.0):
        want = _response(decim, f, fd) * _response(curve, f * R, fd)
        got = _response(folded, f, fd)
        assert abs(got - want) < 2e-6, f
    assert abs(float(np.sum(folded.astype(np.float64))) - 1.0) < 1e-5


def _fm_tone_amplitude(p, modulation_hz):
    """Recovered audio amplitude of an FM tone at the plan's peak deviation,
    through the oracle demodulator (a flat receiver recovers +/-1)."""
    pl = p.plan
    fd = float(pl.demod_rate)
    cap = 1 << 16
    m = np.arange(cap)
    beta = pl.deviation / modulation_hz
    ring = np.exp(1j * beta * np.sin(2 * np.pi * modulation_hz * m / fd)).astype(np.complex64)
    R = pl.demod.decimation
    first = (pl.demod.audio_taps + R) // R + 64          # every tap reads signal
    count = (cap - first * R) // R - 1
    dp = p.demod_block(cap - 1, first, count)
    audio = oracle.vrx_demod(p.demod, dp, ring, p.weights).astype(np.float64)
    t = (first + np.arange(count)) / float(pl.output_rate)
    basis = np.stack([np.cos(2 * np.pi * modulation_hz * t), np.sin(2 * np.pi * modulation_hz * t),
                      np.ones_like(t)], axis=1)
    coef, *_ = np.linalg.lstsq(basis, audio, rcond=None)
    fit = basis @ coef
    purity = 1.0 - np.sum((audio - fit) ** 2) / np.sum((audio - audio.mean()) ** 2)
    return float(np.hypot(coef[0], coef[1])), float(purity)


@pytest.mark.parametrize("curve", [DEEMPH_US75, DEEMPH_EU50])
def test_a_deemphasised_receiver_follows_the_one_pole(curve):
    on = _wfm(deemphasis=curve)
    flat = _wfm(deemphasis=DEEMPH_NONE)
    for f in (500.0, 1000.0, 2122.0, 5000.0, 8000.0):
        amp, purity = _fm_tone_amplitude(on, f)
        ref, _ = _fm_tone_amplitude(flat, f)
        assert ref == pytest.approx(1.0, abs=0.02), f
        err = 20 * np.log10(amp / _analog(curve, f))
        assert abs(err) < WARP_BUDGET_DB, (curve, f, amp, err)
        assert purity > 0.95
