"""The grid vocabulary: channel_count_for (core/engine/vrx_place.cpp:187),
channel_centre (core/dsp/pfb_design.cpp:349) and channel_sample_to_input
(:593), through the C ABI.

Pinned by the reference's own expectations (tests/engine/test_vrx_tune.cpp:
154-176), by exact rational closed forms, and by the group delay's physical
meaning: an impulse at the input index channel_sample_to_input names comes
out of the float64 channelizer definition centred on that channel sample.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from revenant_amd import dsp


def test_channel_count_follows_the_widest_receiver():
    # test_vrx_tune.cpp:154-171 at 2.4 MS/s
    assert dsp.channel_count_for(2_400_000, 200_000) == 8
    assert 2_400_000 // 8 >= 200_000
    assert dsp.channel_count_for(2_400_000, 16_000) == 128
    assert dsp.channel_count_for(250_000, 200_000) == 2
    assert dsp.channel_count_for(0, 200_000) == 2
    assert dsp.channel_count_for(2_400_000, 0) == 2
    assert dsp.channel_count_for(2_400_000, -5) == 2


@pytest.mark.parametrize("rate", [2_400_000, 122_880_000, 20_000_000, 3_200_000, 1_000_003])
@pytest.mark.parametrize("width", [500, 2_700, 10_000, 12_500, 200_000, 999_999])
def test_channel_count_is_the_largest_power_of_two_that_carries(rate, width):
    m = dsp.channel_count_for(rate, width)
    assert m >= 2 and m & (m - 1) == 0
    if m > 2:
        assert rate // m >= width, "the channel must carry the receiver"
    assert rate // (2 * m) < width or 2 * m > 2**31, "twice as many channels would still carry it"


def test_wide_receivers_at_the_wideband_rate_take_512_channels():
    # a 200 kHz receiver at 122.88 MS/s: rate/M >= 200 kHz -> M = 512
    assert dsp.channel_count_for(122_880_000, 200_000) == 512


@pytest.mark.parametrize("M", [2, 8, 64, 4096])
@pytest.mark.parametrize("rate", [2_400_000, 122_880_000, 1_000_003])
def test_channel_centres_are_exact_reduced_rationals(M, rate):
    grid = dsp.GridParams(M, 17, max(1, M // 2))
    for k in list(range(min(M, 70))) + [M // 2, M - 1]:
        num, den = dsp.channel_centre(grid, rate, k)
        signed_k = k if k < M // 2 else k - M
        assert Fraction(num, den) == Fraction(signed_k * rate, M)
        assert den > 0 and np.gcd(num, den) == 1 or (num == 0 and den == 1)
    assert dsp.channel_centre(grid, rate, 0) == (0, 1)
    assert Fraction(*dsp.channel_centre(grid, rate, M - 1)) == Fraction(-rate, M)
    assert Fraction(*dsp.channel_centre(grid, rate, M // 2)) == Fraction(-rate, 2)
    # the channel index is taken modulo M
    assert dsp.channel_centre(grid, rate, M + 1) == dsp.channel_centre(grid, rate, 1)


def test_empty_grid_centre_is_zero():
    assert dsp.channel_centre(dsp.GridParams(0, 17, 1), 2_400_000, 5) == (0, 1)


def test_placement_reports_the_same_centre():
    grid = dsp.GridParams(64, 17, 32)
    for hz in (0, 37_500, 200_000, -200_000, 1_199_000, -1_199_999):
        pl = dsp.place(grid, 2_400_000, dsp.VrxRequest(center_hz=hz, mode=1))
        assert (pl.centre_num, pl.centre_den) == dsp.channel_centre(grid, 2_400_000, pl.channel)


def test_channel_sample_to_input_closed_form_and_wrap():
    grid = dsp.GridParams(64, 17, 32)
    proto = dsp.design_prototype(grid)
    assert proto.group_delay_samples == 64 * 16 // 2
    assert dsp.channel_sample_to_input(grid, proto, 1000, 40) == 1000 + 40 * 32 - 512
    # before the filter filled: wraps near 2^64 instead of clamping
    assert dsp.channel_sample_to_input(grid, proto, 0, 3) == (3 * 32 - 512) % 2**64


def test_an_impulse_comes_out_at_the_channel_sample_that_names_it():
    from oracle import oracle

    grid = dsp.GridParams(8, 17, 4)
    proto = dsp.design_prototype(grid, beta=8.0)
    m0 = 60
    n0 = dsp.channel_sample_to_input(grid, proto, 0, m0)
    assert n0 == m0 * 4 - 8 * 16 // 2
    samples = 4 * 100
    iq = np.zeros(2 * samples, dtype=np.int16)
    iq[2 * n0] = 16384
    first = (grid.prototype_length - 1 + 3) // 4
    y = oracle.fp64_chain(iq.view(np.uint32), proto.taps, 8, 4, first, 100 - first)
    mag = np.abs(y[:, 0])
    peak = int(np.argmax(mag)) + first
    assert peak == m0
    for j in range(1, 8):  # linear phase: symmetric about the centre sample
        assert mag[m0 - first - j] == pytest.approx(mag[m0 - first + j], rel=1e-9, abs=1e-15)
