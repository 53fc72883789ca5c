"""The reference's own planner expectations (tests/reference/test_vrx.cpp:417-1066),
through the C ABI: the passband shorthand, the default passbands, the fit of
each edge into one channel, refusals of what a channel cannot carry, the cheap
demodulation-rate query against the plan, and the minimum-rate closed forms.

The reference's grid throughout: 64 channels, 17 taps, D = 32 at 2.4 MS/s
(75 kS/s channels), and 16 channels, D = 8 for wideband FM.
"""
from __future__ import annotations

import pytest

from revenant_amd import _lib, dsp
from revenant_amd.dsp import (DEMOD_AM, DEMOD_CW, DEMOD_DSB, DEMOD_LSB, DEMOD_NFM, DEMOD_RAW,
                              DEMOD_SAM, DEMOD_USB, DEMOD_WFM, GridParams, VrxRequest)

RATE = 2_400_000
GRID = GridParams(64, 17, 32)
WIDE = GridParams(16, 17, 8)
MODES = (DEMOD_RAW, DEMOD_AM, DEMOD_NFM, DEMOD_WFM, DEMOD_USB, DEMOD_LSB, DEMOD_DSB, DEMOD_CW)


def test_the_bandwidth_shorthand_expands_per_mode():
    # test_vrx.cpp:417-480
    want = {DEMOD_RAW: (-1500, 1500), DEMOD_AM: (-1500, 1500), DEMOD_NFM: (-1500, 1500),
            DEMOD_WFM: (-1500, 1500), DEMOD_USB: (0, 3000), DEMOD_LSB: (-3000, 0),
            DEMOD_DSB: (-1500, 1500), DEMOD_CW: (-1500, 1500)}
    for mode, band in want.items():
        got = dsp.resolve_passband(VrxRequest(0, mode, 3000))
        assert got == band and got[1] - got[0] == 3000, mode
    # an odd width resolves symmetric and one hertz narrow
    assert dsp.resolve_passband(VrxRequest(0, DEMOD_NFM, 501)) == (-250, 250)
    # the pair wins over the shorthand
    assert dsp.resolve_passband(VrxRequest(0, DEMOD_USB, 99_999, passband=(300, 2700))) == \
        (300, 2700)
    # an inverted band is refused with both numbers named
    with pytest.raises(_lib.RvxError) as e:
        dsp.resolve_passband(VrxRequest(0, DEMOD_NFM, passband=(2700, 300)))
    assert "2700" in str(e.value) and "300" in str(e.value)


def test_every_mode_has_a_default_passband_and_ssb_excludes_the_carrier():
    # test_vrx.cpp:482-516
    want = {DEMOD_RAW: (-6000, 6000), DEMOD_AM: (-5000, 5000), DEMOD_NFM: (-8000, 8000),
            DEMOD_WFM: (-100_000, 100_000), DEMOD_USB: (300, 2700), DEMOD_LSB: (-2700, -300),
            DEMOD_DSB: (-3000, 3000), DEMOD_CW: (-600, 600)}
    for mode, band in want.items():
        assert dsp.default_passband(mode) == band, mode
        assert dsp.resolve_passband(VrxRequest(0, mode)) == band
    usb, lsb = dsp.default_passband(DEMOD_USB), dsp.default_passband(DEMOD_LSB)
    assert usb[0] > 0 and lsb[1] < 0 and usb == (-lsb[1], -lsb[0])


def _usb_placed():
    return dsp.place(GRID, RATE, VrxRequest(196_500, DEMOD_USB, passband=(300, 2700)))


def test_the_fit_moves_only_the_edge_that_does_not_fit():
    # test_vrx.cpp:518-572
    pl = _usb_placed()
    assert (pl.granted_low, pl.granted_high, pl.bandwidth_clamped) == (300, 2700, 0)
    limit = dsp.max_channel_bandwidth(pl) // 2
    assert limit > 0
    assert dsp.clamp_to_channel(pl, (-1000, limit + 5000)) == (-1000, limit)
    assert dsp.clamp_to_channel(pl, (-limit, limit)) == (-limit, limit)
    assert dsp.clamp_to_channel(pl, (-limit - 5000, 1000)) == (-limit, 1000)
    assert dsp.clamp_to_channel(pl, (-limit - 5000, limit + 5000)) == (-limit, limit)
    # place reports the same fit: one edge moved, the other kept, clamped flagged
    high = dsp.place(GRID, RATE, VrxRequest(196_500, DEMOD_USB, passband=(-1000, limit + 5000)))
    assert (high.granted_low, high.granted_high, high.bandwidth_clamped) == (-1000, limit, 1)
    # the limit is the channel rate less twice |residual|, halved, rounded down
    res = abs(pl.residual_num) / pl.residual_den
    assert dsp.max_channel_bandwidth(pl) == int((pl.channel_rate * pl.residual_den -
                                                  2 * abs(pl.residual_num)) // pl.residual_den)
    assert dsp.max_channel_bandwidth(pl) <= pl.channel_rate - 2 * res


def test_a_band_the_channel_cannot_reach_comes_back_empty_and_is_refused():
    # test_vrx.cpp:574-620
    pl = _usb_placed()
    limit = dsp.max_channel_bandwidth(pl) // 2
    lo, hi = dsp.clamp_to_channel(pl, (limit + 1000, limit + 2000))
    assert lo <= hi and hi - lo == 0
    lo, hi = dsp.clamp_to_channel(pl, (-limit - 2000, -limit - 1000))
    assert lo <= hi and hi - lo == 0
    outside = VrxRequest(196_500, DEMOD_USB, passband=(limit + 1000, limit + 2000))
    with pytest.raises(_lib.RvxError) as e:
        dsp.plan_vrx(GRID, RATE, outside, pl)
    assert str(limit + 1000) in str(e.value) and str(limit) in str(e.value)
    with pytest.raises(_lib.RvxError):
        dsp.demod_rate_for(outside, pl, 48_000)


def test_a_placement_that_carries_nothing_is_refused_not_planned():
    # test_vrx.cpp:622-656: built by hand, half a channel off its centre
    broken = _lib.VrxPlacement()
    broken.channel, broken.channel_rate = 3, 75_000
    broken.residual_num, broken.residual_den = 40_000, 1
    broken.centre_num, broken.centre_den = 112_500, 1
    assert dsp.max_channel_bandwidth(broken) == 0
    assert dsp.clamp_to_channel(broken, (-4000, 4000)) == (0, 0)
    req = VrxRequest(196_500, DEMOD_NFM, passband=(-4000, 4000))
    with pytest.raises(_lib.RvxError) as e:
        dsp.plan_vrx(GRID, RATE, req, broken)
    assert "nothing" in str(e.value)
    with pytest.raises(_lib.RvxError):
        dsp.demod_rate_for(req, broken, 48_000)


@pytest.mark.parametrize("mode,low,high,audio_rate", [
    (DEMOD_RAW, -6000, 6000, 48_000), (DEMOD_AM, -5000, 5000, 48_000),
    (DEMOD_NFM, -8000, 8000, 48_000), (DEMOD_USB, 300, 2700, 48_000),
    (DEMOD_LSB, -2700, -300, 16_000), (DEMOD_DSB, -3000, 3000, 24_000),
    (DEMOD_CW, -250, 250, 8_000),
    # off-centre and asymmetric: the reach term decides
    (DEMOD_NFM, 1000, 9000, 16_000), (DEMOD_USB, -9000, -1000, 12_000),
    # SAM at an audio rate low enough that its own minimum decides
    (DEMOD_SAM, -5000, 5000, 8_000)])
def test_the_cheap_rate_query_answers_what_the_planner_builds(mode, low, high, audio_rate):
    # test_vrx.cpp:658-725 (the digital voice modes are not planned by this build)
    req = VrxRequest(196_500, mode, passband=(low, high), audio_rate=audio_rate)
    pl = dsp.place(GRID, RATE, req)
    cheap = dsp.demod_rate_for(req, pl, audio_rate)
    plan = dsp.plan_vrx(GRID, RATE, req, pl)
    assert cheap == plan.plan.demod_rate
    assert cheap % audio_rate == 0


def _retired_minimum(mode, bandwidth, pitch):
    """test_vrx.cpp:727-752: the one-width function the general rule replaced."""
    if bandwidth <= 0:
        return 0
    floor_rate = (3 * bandwidth + 1) // 2
    if mode in (DEMOD_AM, DEMOD_USB, DEMOD_LSB):
        return max(floor_rate, 2 * bandwidth)
    if mode == DEMOD_CW:
        return max(floor_rate, 2 * pitch + bandwidth)
    return floor_rate


def test_the_minimum_rate_against_its_closed_form_and_the_retired_shorthand():
    # test_vrx.cpp:754-850, every cell
    for pitch in (0, 1, 250, 700, 5000):
        for bw in (1, 2, 3, 500, 501, 3000, 3001, 12_000, 12_001):
            half, even = bw // 2, bw % 2 == 0
            for mode in MODES:
                now = dsp.minimum_demod_rate(mode, bw, pitch)
                before = _retired_minimum(mode, bw, pitch)
                if mode in (DEMOD_USB, DEMOD_LSB):
                    expect, g_lo, g_hi = 2 * bw, 0, 0
                elif mode == DEMOD_AM:
                    expect = 4 * half
                    g_lo = g_hi = 0 if even else 2
                elif mode == DEMOD_CW:
                    expect = max(3 * half, 2 * pitch + 2 * half)
                    g_lo, g_hi = (0, 0) if even else (1, 2)
                else:
                    expect = 3 * half
                    g_lo = g_hi = 0 if even else 2
                if bw == 1 and mode not in (DEMOD_USB, DEMOD_LSB):
                    expect = 0
                    g_lo = g_hi = max(2, 2 * pitch + 1) if mode == DEMOD_CW else 2
                assert now == expect, (mode, bw, pitch)
                assert g_lo <= before - now <= g_hi, (mode, bw, pitch, before, now)
    # a one-hertz symmetric request: place grants an empty band, the planner refuses
    narrow = VrxRequest(196_500, DEMOD_NFM, 1)
    pl = dsp.place(GRID, RATE, narrow)
    assert pl.granted_high == pl.granted_low
    with pytest.raises(_lib.RvxError):
        dsp.plan_vrx(GRID, RATE, narrow, pl)
    with pytest.raises(_lib.RvxError):
        dsp.demod_rate_for(narrow, pl, 48_000)
    # a negative pitch reads as zero
    for bw in (1000, 3000):
        assert dsp.minimum_demod_rate(DEMOD_CW, bw, -3000) == dsp.minimum_demod_rate(DEMOD_CW, bw, 0)


@pytest.mark.parametrize("mode,bw,grid", [
    (DEMOD_RAW, 12_000, GRID), (DEMOD_RAW, 2400, GRID), (DEMOD_AM, 10_000, GRID),
    (DEMOD_AM, 6000, GRID), (DEMOD_NFM, 12_000, GRID), (DEMOD_NFM, 25_000, GRID),
    (DEMOD_WFM, 200_000, WIDE), (DEMOD_WFM, 150_000, WIDE), (DEMOD_USB, 2700, GRID),
    (DEMOD_USB, 1800, GRID), (DEMOD_LSB, 2700, GRID), (DEMOD_LSB, 1800, GRID),
    (DEMOD_DSB, 6000, GRID), (DEMOD_DSB, 3000, GRID), (DEMOD_CW, 500, GRID),
    (DEMOD_CW, 250, GRID), (DEMOD_SAM, 10_000, GRID), (DEMOD_SAM, 6000, GRID)])
def test_a_plans_fine_table_is_the_length_its_config_implies(mode, bw, grid):
    # test_vrx.cpp:859-905
    centre = 160_000 if grid is WIDE else 196_500
    plan = dsp.plan_vrx(grid, RATE, VrxRequest(centre, mode, bw))
    assert plan.fine_taps.size == plan.fine.phases * plan.fine.taps + 1


def test_the_raw_tap_is_a_complex_pair_with_no_filter():
    # the raw tap's shape (test_vrx.cpp:990-1012, as this build plans the raw mode)
    plan = dsp.plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_RAW))
    d = plan.demod
    assert (d.channels, d.decimation, d.audio_taps, d.dc_taps, d.pilot_taps) == (2, 1, 1, 1, 0)
    assert plan.plan.demod_gain == 1.0 and plan.plan.deviation == 0
    assert plan.plan.deemphasis == dsp.DEEMPH_NONE and not plan.plan.stereo
    assert plan.history == 0
    assert (plan.plan.passband_low, plan.plan.passband_high) == dsp.default_passband(DEMOD_RAW)
    assert not plan.plan.placement.bandwidth_clamped


def test_an_fm_receiver_its_channel_cannot_carry_is_refused_with_the_count_that_would():
    # vrx_place.cpp:398-440: a 200 kHz WFM receiver on the 75 kS/s grid
    with pytest.raises(_lib.RvxError) as e:
        dsp.place(GRID, RATE, VrxRequest(196_500, DEMOD_WFM, 200_000))
    assert "8 channels" in str(e.value)
    # an AM receiver clamped the same way is narrowed, not refused
    pl = dsp.place(GRID, RATE, VrxRequest(196_500, DEMOD_AM, 200_000))
    assert pl.bandwidth_clamped and pl.granted_high - pl.granted_low < 200_000
