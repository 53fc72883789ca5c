"""Comb detection (rvx_comb_detect; core/dsp/comb_detect.cpp) on spectra whose
comb is chosen, so every assertion is against arithmetic: the cases of
tests/detect/test_comb_detect.cpp:149-644, seeded as there (fewer seeds where
the reference runs hundreds of pure-noise scenes)."""
import math

import numpy as np
import pytest

from comb_scenes import BINS, HZ_PER_BIN, RATE, Scene, line_nearest_dc, on_comb
from revenant_amd import _lib
from revenant_amd.comb import detect_comb, detect_combs


def check(found, c, d):
    """Spacing to 1 Hz, the k = 0 line to a tenth of a bin (test_comb_detect.cpp:138)."""
    assert found is not None
    assert abs(found.spacing_hz - d * HZ_PER_BIN) < 1.0
    assert abs(found.offset_hz - line_nearest_dc(c, d)) < 0.1 * HZ_PER_BIN
    assert abs(found.offset_hz) <= 0.5 * found.spacing_hz + 1.0
    assert found.k_lo < 0 < found.k_hi


def line_count(e):
    return 0 if e.k_hi < e.k_lo else e.k_hi - e.k_lo + 1


@pytest.mark.parametrize("seed", range(1, 9))
def test_recovers_c_and_d_of_an_hf_switching_comb(seed):
    s = Scene(seed)
    d = 14_000.0 / HZ_PER_BIN          # 95.57 bins
    c = 37.31 + 0.11 * seed
    drawn = s.comb(c, d, 4.0, 15.0)
    found = detect_comb(s.render(), RATE)
    check(found, c, d)
    assert found.lines_found >= drawn * 8 // 10
    assert line_count(found) <= drawn
    assert 6.0 < found.median_height_db < 13.0


@pytest.mark.parametrize("seed", range(100, 108))
def test_keeps_one_comb_across_gaps(seed):
    s = Scene(seed)
    s.comb(12.6, 61.37, 6.0, 15.0, 0.6)
    found = detect_comb(s.render(), RATE)
    check(found, 12.6, 61.37)
    assert found.lines_found < line_count(found)


@pytest.mark.parametrize("seed", range(200, 208))
def test_is_not_fooled_by_carriers_and_wideband_signals(seed):
    s = Scene(seed)
    d = 14_000.0 / HZ_PER_BIN
    c = 50.42
    s.comb(c, d, 4.0, 15.0, 0.9)
    for _ in range(30):  # broadcast carriers much stronger than the comb
        s.line(s.g.uniform(10.0, BINS - 10.0), s.g.uniform(20.0, 35.0))
    for k in (3, 40, 77, 150):  # four sitting exactly on the grid
        s.line(c + k * d, 30.0)
    for i in range(8):  # stations 10 kHz wide, 10 to 25 dB up
        s.plateau(int(s.g.integers(0, BINS - 80)), 68, 10.0 + 2.0 * i)
    check(detect_comb(s.render(), RATE), c, d)


@pytest.mark.parametrize("seed", range(300, 306))
def test_alternate_lines_weak_picks_d_not_2d(seed):
    s = Scene(seed)
    d, c = 40.25, 7.7
    for k in range(-300, 300):
        x = c + k * d + 8192.0
        if 3.0 < x < BINS - 3.0:
            s.line(x, 15.0 if k % 2 == 0 else 6.0)
    check(detect_comb(s.render(), RATE), c + 8192.0, d)


@pytest.mark.parametrize("seed", range(310, 316))
def test_strays_at_the_midpoints_pick_d_not_half_d(seed):
    s = Scene(seed)
    d, c = 80.6, 31.2
    s.comb(c, d, 6.0, 15.0)
    for k in (5, 30, 61, 90, 120, 150, 180):
        s.line(c + (k + 0.5) * d, 10.0)
    check(detect_comb(s.render(), RATE), c, d)


@pytest.mark.parametrize("seed", range(400, 404))
def test_resolves_a_comb_four_bins_apart_and_refuses_one_closer(seed):
    s = Scene(seed)
    s.comb(1.3, 4.6, 8.0, 15.0)
    check(detect_comb(s.render(), RATE), 1.3, 4.6)
    s = Scene(seed)
    s.comb(0.9, 2.5, 8.0, 15.0)
    assert detect_comb(s.render(), RATE) is None


@pytest.mark.parametrize("seed", range(500, 520))
def test_finds_nothing_in_carriers_and_stations(seed):
    s = Scene(seed)
    for _ in range(40):
        s.line(s.g.uniform(10.0, BINS - 10.0), s.g.uniform(5.0, 35.0))
    for i in range(10):
        s.plateau(int(s.g.integers(0, BINS - 80)), 68, 8.0 + 2.0 * i)
    assert detect_comb(s.render(), RATE) is None


def test_finds_nothing_in_averaged_noise_or_one_frame_or_zeros():
    for seed in range(1000, 1060):
        assert detect_comb(Scene(seed).render(), RATE) is None, seed
    for seed in range(2000, 2050):
        assert detect_comb(Scene(seed, frames=1.0).render(), RATE) is None, seed
    assert detect_comb(np.zeros(BINS, np.float32), RATE) is None
    assert detect_combs(Scene(600).render(), RATE) == []
    assert detect_combs(np.zeros(BINS, np.float32), RATE) == []


def test_refuses_a_spectrum_it_cannot_read():
    ok = np.ones(BINS, np.float32)
    for power, rate in ((np.zeros(0, np.float32), RATE), (ok, 0), (ok, -5),
                        (np.ones(32, np.float32), RATE)):
        with pytest.raises(_lib.RvxError):
            detect_comb(power, rate)
    for bad in (math.nan, -1.0, math.inf):
        p = ok.copy()
        p[17] = bad
        with pytest.raises(_lib.RvxError):
            detect_comb(p, RATE)
        with pytest.raises(_lib.RvxError):
            detect_combs(p, RATE)


@pytest.mark.parametrize("seed", range(300, 308))
def test_separates_two_interleaved_switchers(seed):
    """Two switchers near 31.58 kHz a few hertz apart in spacing, interleaved
    to look like one comb at 15.79 kHz: each reported at its own spacing."""
    s = Scene(seed)
    d_a, d_b = 31_574.0 / HZ_PER_BIN, 31_579.5 / HZ_PER_BIN
    c_a = 41.3 + 0.07 * (seed - 300)
    c_b = c_a + 15_790.0 / HZ_PER_BIN
    s.comb(c_a, d_a, 9.0, 14.0)
    s.comb(c_b, d_b, 5.0, 9.0)
    found = detect_combs(s.render(), RATE)
    assert len(found) >= 2
    assert on_comb(found[0], c_a, d_a, 2.0) and on_comb(found[1], c_b, d_b, 2.0)
    assert found[0].lines_found >= 50 and found[1].lines_found >= 50


@pytest.mark.parametrize("seed", range(400, 408))
def test_holds_a_single_comb_as_its_two_exact_halves(seed):
    s = Scene(seed)
    d = 14_000.0 / HZ_PER_BIN
    c = 37.31 + 0.11 * (seed - 400)
    drawn = s.comb(c, d, 6.0, 15.0)
    found = detect_combs(s.render(), RATE)
    assert len(found) == 2
    even, odd = (found[0], found[1]) if on_comb(found[0], c, 2 * d, 1.0) else (found[1], found[0])
    assert on_comb(even, c, 2 * d, 1.0) and on_comb(odd, c + d, 2 * d, 1.0)
    assert even.lines_found + odd.lines_found >= drawn * 8 // 10


@pytest.mark.parametrize("seed", range(800, 804))
def test_reports_a_combs_ripple_sidebands_with_it(seed):
    s = Scene(seed)
    d = 31_578.9 / HZ_PER_BIN
    c = 45.2 + 0.13 * (seed - 800)
    ripple = 3.4  # bins
    k_lo, k_hi = int(np.ceil((20.0 - c) / d)), int(np.floor((BINS - 20.0 - c) / d))
    for k in range(k_lo, k_hi + 1):
        x = c + k * d
        s.line(x, 22.0)
        for m, level in ((1, -12.0), (2, -16.0)):
            s.line(x + m * ripple, 22.0 + level)
            s.line(x - m * ripple, 22.0 + level)
    found = detect_combs(s.render(), RATE)
    assert found
    for f in found:
        assert on_comb(f, c, d, 2.0) or on_comb(f, c, 2 * d, 2.0) or on_comb(f, c + d, 2 * d, 2.0)
        assert abs(f.ripple_hz - ripple * HZ_PER_BIN) < 0.3 * HZ_PER_BIN
        assert f.sideband_orders > 0


@pytest.mark.parametrize("seed", range(700, 706))
def test_separates_interleaved_combs_offset_from_half_a_spacing(seed):
    s = Scene(seed)
    d = 31_578.9 / HZ_PER_BIN
    c_a = 45.2 + 0.13 * (seed - 700)
    c_b = c_a + (15_789.45 + 100.0) / HZ_PER_BIN
    s.comb(c_a, d, 9.0, 14.0)
    s.comb(c_b, d, 8.0, 13.0)
    found = detect_combs(s.render(), RATE)
    assert any(on_comb(f, c_a, d, 2.0) for f in found)
    assert any(on_comb(f, c_b, d, 2.0) for f in found)


def _held(found, c, d):
    whole = any(on_comb(f, c, d, 1.0) for f in found)
    even = any(on_comb(f, c, 2 * d, 1.0) for f in found)
    odd = any(on_comb(f, c + d, 2 * d, 1.0) for f in found)
    return whole or (even and odd)


@pytest.mark.parametrize("seed", range(500, 504))
def test_finds_an_unrelated_comb_beside_the_first(seed):
    s = Scene(seed)
    d_a, d_b = 14_000.0 / HZ_PER_BIN, 22_930.0 / HZ_PER_BIN
    s.comb(37.31, d_a, 10.0, 15.0)
    s.comb(80.9, d_b, 6.0, 9.0)
    found = detect_combs(s.render(), RATE)
    assert _held(found, 37.31, d_a) and _held(found, 80.9, d_b)
    assert len(found) <= _lib.COMB_MAX_COMBS


def _held_comb(c, d):
    """A comb at c + k d bins over the span, as the bank reports a tracker holding it."""
    e = _lib.CombEstimate()
    e.offset_hz = line_nearest_dc(c, d)
    e.spacing_hz = d * HZ_PER_BIN
    c0 = BINS // 2 + e.offset_hz / HZ_PER_BIN
    e.k_lo = int(np.ceil((2.0 - c0) / d))
    e.k_hi = int(np.floor((BINS - 3.0 - c0) / d))
    return e


@pytest.mark.parametrize("seed", range(900, 906))
def test_skips_a_comb_already_held_and_finds_the_fresh_one(seed):
    s = Scene(seed)
    d_a, c_a = 14_000.0 / HZ_PER_BIN, 37.31 + 0.05 * (seed - 900)
    d_b, c_b = 22_930.0 / HZ_PER_BIN, 80.9
    s.comb(c_a, d_a, 12.0, 18.0)
    s.comb(c_b, d_b, 5.0, 8.0)
    found = detect_combs(s.render(), RATE, active=[_held_comb(c_a, d_a)])
    assert found
    for f in found:
        assert not on_comb(f, c_a, d_a, 2.0)
        assert not on_comb(f, c_a, 2 * d_a, 2.0) and not on_comb(f, c_a + d_a, 2 * d_a, 2.0)
    assert _held(found, c_b, d_b)


@pytest.mark.parametrize("seed", range(910, 914))
def test_no_held_comb_changes_nothing(seed):
    s = Scene(seed)
    s.comb(37.31, 14_000.0 / HZ_PER_BIN, 10.0, 15.0)
    s.comb(80.9, 22_930.0 / HZ_PER_BIN, 6.0, 9.0)
    p = s.render()
    plain, empty = detect_combs(p, RATE), detect_combs(p, RATE, active=[])
    assert len(plain) == len(empty)
    for a, b in zip(plain, empty):
        assert (a.offset_hz, a.spacing_hz, a.lines_found) == (b.offset_hz, b.spacing_hz,
                                                              b.lines_found)
