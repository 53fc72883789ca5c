"""Hybrid digital sidebands beside wideband FM (core/detect/hd_hybrid.cpp,
Detector::mark_hd_hybrids detector.cpp:1943-1976), pinned on constructed
spectra whose readings have closed forms: flat sidebands of excess S over a
unit floor sit 10 log10(1 + S) over their floor and over an empty gap, their
level against a flat analogue excess A over 201 one-kHz bins is
10 log10(70 S / (201 A)) dBc (70 bins a side), and a side whose inner half
carries a and outer half b is tilted by 10 log10((1 + a) / (1 + b)). A reading
fails on a weak side, a filled gap, a tilt over 3 dB or sides more than 6 dB
apart; a track is confirmed after two passing decisions, its levels taken whole
then smoothed by 0.3, and released after three failing ones. Host only."""
import math

import numpy as np
import pytest

from revenant_amd import _lib, events

BINS = 4096
RATE = 4_096_000
BW = RATE / BINS                    # 1 kHz bins
ZERO = -BINS / 2 * BW
CENTER = 100_000_000
C_BIN = 1500                        # the station's centre bin
C_HZ = ZERO + C_BIN * BW            # relative to the source centre


def spectrum(analogue=2.0, lower=(3.0, 3.0), upper=(3.0, 3.0), gap=0.0):
    """Unit floor; analogue excess over +-100 kHz; each side's inner
    (129-163 kHz) and outer (164-198 kHz) excess; the 105-120 kHz gaps."""
    power, floor = np.ones(BINS), np.ones(BINS)
    off = np.arange(BINS) - C_BIN     # kHz from the centre
    power[np.abs(off) <= 100] += analogue
    for sign, (inner, outer) in ((-1, lower), (1, upper)):
        k = sign * off
        power[(k >= 129) & (k <= 163)] += inner
        power[(k >= 164) & (k <= 198)] += outer
        power[(k >= 105) & (k <= 120)] += gap
    return power, floor


def read(power, floor, centre=C_HZ):
    return events.measure_hd_sidebands(power, floor, ZERO, BW, centre)


def db(x):
    return 10.0 * math.log10(x)


def test_flat_sidebands_read_their_closed_forms():
    for s_level, a in ((3.0, 2.0), (1.0, 10.0), (0.5, 1.0)):
        r = read(*spectrum(analogue=a, lower=(s_level, s_level), upper=(s_level, s_level)))
        assert r.measured
        assert r.centre_hz == pytest.approx(C_HZ, abs=1e-6)
        assert r.lower_dbc == pytest.approx(db(70 * s_level / (201 * a)), abs=1e-12)
        assert r.upper_dbc == pytest.approx(r.lower_dbc, abs=1e-12)
        for v in (r.lower_over_floor_db, r.upper_over_floor_db, r.lower_over_gap_db,
                  r.upper_over_gap_db):
            assert v == pytest.approx(db(1 + s_level), abs=1e-12)
        assert abs(r.lower_tilt_db) < 1e-12 and abs(r.upper_tilt_db) < 1e-12
        # 3 dB over the floor needs S >= 10^0.3 - 1 = 0.995
        assert bool(r.hybrid) == (db(1 + s_level) >= 3.0)


def test_the_centroid_finds_the_analogue_signal():
    # asked 30 kHz off, two centroid passes walk most of the way back. Pass 1
    # (offsets -70..+130 kHz) holds the analogue excess 2 at -70..+100 and the
    # upper sideband's 3 at +129, +130: (2 * 2565 + 3 * 259) / (2 * 171 + 6)
    # = 16.97 kHz. Pass 2 (-83..+116) holds the analogue alone, offsets -83..100:
    # mean 8.5 kHz, exactly.
    r = read(*spectrum(), centre=C_HZ + 30_000)
    assert r.centre_hz == pytest.approx(C_HZ + 8_500, abs=1e-6)
    assert r.measured and r.hybrid


@pytest.mark.parametrize("case", ["tilt", "asymmetry", "weak_side", "gap_filled", "edge"])
def test_what_is_not_a_hybrid(case):
    if case == "tilt":                # inner 4, outer 1: tilt 10 log10(5/2) = 3.98 dB
        r = read(*spectrum(lower=(4.0, 1.0)))
        assert r.lower_tilt_db == pytest.approx(db(5 / 2), abs=1e-12)
    elif case == "asymmetry":         # 9 against 1: 9.54 dB apart
        r = read(*spectrum(lower=(1.0, 1.0), upper=(9.0, 9.0)))
        assert r.upper_dbc - r.lower_dbc == pytest.approx(db(9.0), abs=1e-12)
    elif case == "weak_side":
        r = read(*spectrum(upper=(0.5, 0.5)))
        assert r.upper_over_floor_db == pytest.approx(db(1.5), abs=1e-12)
    elif case == "gap_filled":        # the gap as full as the sidebands: 0 dB over it
        r = read(*spectrum(gap=3.0))
        assert r.lower_over_gap_db == pytest.approx(0.0, abs=1e-12)
    else:                             # the outer sideband off the spectrum: not measured
        r = read(*spectrum(), centre=ZERO + 150 * BW)
        assert not r.measured
    assert not r.hybrid


def _track(t, k, power, floor):
    c = _lib.Candidate()
    c.first_bin, c.last_bin = C_BIN - 100, C_BIN + 100
    c.center_hz, c.bandwidth_hz = CENTER + int(C_HZ), 201_000
    c.snr_2500_db = 30.0
    t.update(power, floor, [c], 1_000 + k * RATE // 10, 0.1)
    return t.tracks()[0]


def test_a_track_is_confirmed_after_two_and_released_after_three():
    t = events.Tracker(events.event_config(BW, ZERO, RATE, CENTER, residual_decisions=0,
                                           centre_smoothing=1.0), BINS)
    hybrid, plain = spectrum(lower=(3.0, 3.0), upper=(3.0, 3.0)), spectrum(lower=(0.0, 0.0),
                                                                           upper=(0.0, 0.0))
    k = 0
    for _ in range(2):                # pending: not measured
        tr = _track(t, k, *hybrid); k += 1
        assert tr.state == 0 and tr.hd_passes == 0
    tr = _track(t, k, *hybrid); k += 1  # live: the first pass, levels taken whole
    first = db(70 * 3.0 / (201 * 2.0))
    assert tr.hd_passes == 1 and not tr.hd_hybrid
    assert tr.hd_lower_dbc == pytest.approx(first, abs=1e-12)
    stronger = spectrum(lower=(6.0, 6.0), upper=(6.0, 6.0))
    tr = _track(t, k, *stronger); k += 1  # the second: confirmed, smoothed by 0.3
    assert tr.hd_hybrid and tr.hd_passes == 2
    want = first + 0.3 * (db(70 * 6.0 / (201 * 2.0)) - first)
    assert tr.hd_lower_dbc == pytest.approx(want, abs=1e-12)
    for n in (1, 2):                  # two failing decisions: still held
        tr = _track(t, k, *plain); k += 1
        assert tr.hd_hybrid and tr.hd_fails == n and tr.hd_passes == 0
    tr = _track(t, k, *plain); k += 1   # the third: released
    assert not tr.hd_hybrid and tr.hd_fails == 3


def test_narrow_tracks_are_not_measured():
    t = events.Tracker(events.event_config(BW, ZERO, RATE, CENTER, residual_decisions=0,
                                           centre_smoothing=1.0), BINS)
    power, floor = spectrum()
    for k in range(5):
        c = _lib.Candidate()
        c.first_bin, c.last_bin = C_BIN - 20, C_BIN + 20
        c.center_hz, c.bandwidth_hz = CENTER + int(C_HZ), 41_000   # under 50 kHz
        c.snr_2500_db = 30.0
        t.update(power, floor, [c], 1_000 + k * RATE // 10, 0.1)
    tr = t.tracks()[0]
    assert tr.state == 1 and tr.hd_passes == 0 and not tr.hd_hybrid
