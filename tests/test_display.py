"""The display tap's planning (core/dsp/vrx_reference.cpp:2938-3060), after
tests/reference/test_vrx.cpp:1387-1462: the ladder's figures and properties,
the decimation each reference case gets, the pane holding still through an
edge drag, and the anti-alias table's response measured from its taps (flat
to the pane's edge, in stopband past the fold)."""
import numpy as np
import pytest

from revenant_amd.dsp import (DEMOD_USB, DISPLAY_TAPS, VrxRequest, display_decimation,
                              display_ladder, display_max_decimation, plan_vrx, plan_vrx_display)
from tests.display_cases import GRID, RATE, display_cases


def test_the_ladder_and_its_rungs():
    # 256 taps reach 80 dB across Fdisp / 2 up to R = 25:
    # 2.285 * 2 pi * 255 / (2 R) + 8 >= 80
    assert display_max_decimation() == 25
    assert display_ladder(625_000) == [1, 2, 4, 8, 20]
    assert display_ladder(75_000) == [1, 2, 4, 8, 20]
    assert display_ladder(75_011) == [1]
    for rate in (625_000, 75_000, 300_000, 187_500, 64_000):
        ladder = display_ladder(rate)
        assert ladder[0] == 1
        for i, r in enumerate(ladder):
            assert rate % r == 0 and r <= 25
            if i:
                assert r >= 2 * ladder[i - 1]
        reach = 50
        while reach < rate:
            r = display_decimation(rate, reach)
            pane = rate / r / 2.0
            assert pane >= 4.0 * reach or r == 1
            k = ladder.index(r)
            if k + 1 < len(ladder):
                assert rate / ladder[k + 1] / 2.0 < 4.0 * reach
            reach += reach // 7 + 1


@pytest.mark.parametrize("case", display_cases(), ids=lambda c: c[0])
def test_each_case_gets_the_reference_decimation(case):
    name, plan, want = case
    d = plan_vrx_display(plan)
    p = d.plan
    assert p.decimation == want, name
    assert (p.fine.taps, p.fine.phases) == (DISPLAY_TAPS, 1)
    assert p.display_rate * p.decimation == p.channel_rate
    assert p.group_delay_channel_samples == 127.5
    assert d.taps.size == DISPLAY_TAPS + 1 and d.taps[-1] == 0
    fp, _ = d.block(4096, 4095, 8191, 3_000_000_017, 1024)
    assert (fp.step_rem, fp.frac0, fp.step_whole) == (0, 0, p.decimation)


def test_the_pane_moves_only_when_a_dragged_edge_crosses_a_rung():
    rates = []
    for high in range(1_000, 20_001, 250):
        plan = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_USB, passband=(300, high)))
        r = plan_vrx_display(plan).plan.display_rate
        if not rates or rates[-1] != r:
            rates.append(r)
    assert len(rates) <= len(display_ladder(75_000))
    assert all(b >= 2 * a for a, b in zip(rates, rates[1:]))


@pytest.mark.parametrize("case", display_cases(), ids=lambda c: c[0])
def test_the_table_is_flat_across_the_pane_and_deep_past_the_fold(case):
    """The table demodulated by its centre (the mix), transformed finely:
    within 0.05 dB of unit gain to pass_hz, at least the designed depth less
    two dB beyond stop_hz (checked up to Fc / 2)."""
    _, plan, _ = case
    d = plan_vrx_display(plan)
    p = d.plan
    taps = d.taps[:DISPLAY_TAPS].astype(np.complex128)
    mix = plan.plan.mix_num / plan.plan.mix_den
    k = np.arange(DISPLAY_TAPS) - (DISPLAY_TAPS - 1) / 2.0
    base = taps * np.exp(-2j * np.pi * mix * k / p.channel_rate)
    n = 1 << 16
    h = np.fft.fft(base, n)
    f = np.fft.fftfreq(n, 1.0 / p.channel_rate)
    mag = 20 * np.log10(np.abs(h) + 1e-300)
    inside = np.abs(f) <= p.pass_hz
    assert np.abs(mag[inside]).max() < 0.05
    outside = (np.abs(f) >= p.stop_hz) & (np.abs(f) <= p.channel_rate / 2)
    if outside.any():
        assert mag[outside].max() < -(p.attenuation_db - 2.0)  # Kaiser's estimate is good to ~1 dB
