"""The track fields around the event stages that the oracle does not restate,
pinned by what their definitions fix (core/detect/detector.cpp):
margin confidence (characterise::margin_confidence: 0 under the threshold,
one half at it, 1 - exp(-6/6)/2 six dB over, toward one), the per-frame band
level statistics (measure_levels: Welford mean and squared deviations of the
band's summed linear power in dB, restarted when the band moves, live tracks
only) and the grid channel with its hysteresis (update_channel: nearest
channel, moved only past 0.5 + 0.15 channels)."""
import math

import numpy as np
import pytest

from revenant_amd import events

BINS = 4096
RATE = 122_880_000
M = 4096                      # channels of the grid the tracks are placed on
BW = RATE / BINS              # one bin per channel here, 30 kHz
ZERO = -BINS / 2 * BW
CENTER = 96_000_000


def _candidate(first, last, center_hz, snr_db=20.0):
    c = events._lib.Candidate()
    c.first_bin, c.last_bin = first, last
    c.center_hz, c.bandwidth_hz = int(center_hz), int((last - first + 1) * BW)
    c.snr_2500_db = snr_db
    return c


def _tracker(**kw):
    cfg = events.event_config(BW, ZERO, RATE, CENTER, residual_decisions=0,
                              centre_smoothing=1.0, **kw)
    return events.Tracker(cfg, BINS), cfg


def _step(t, cands, k):
    avg, flo = np.full(BINS, 2.0), np.ones(BINS)
    t.update(avg, flo, cands, 1_000_000 + k * RATE // 10, 0.1)
    return t.tracks()


def test_margin_confidence_of_candidates_and_tracks():
    cfg = events.event_config(BW, ZERO, RATE, CENTER)
    thr = cfg.detection_threshold_db
    g = np.random.default_rng(5)
    avg = g.gamma(64.0, 1.0 / 64.0, BINS)
    floor = np.ones(BINS)
    bands = []
    for start, width, level in ((100, 8, 3.0), (900, 40, 0.6), (2000, 200, 0.08), (3000, 3, 30.0)):
        avg[start:start + width] += level
        bands.append((start, width))
    cands = events.candidates(cfg, avg, floor, [(s, w, 0.0) for s, w in bands])
    assert len(cands) >= 3
    for c in cands:
        want = 0.0 if c.snr_2500_db < thr else 1.0 - 0.5 * math.exp(-(c.snr_2500_db - thr) / 6.0)
        assert c.margin_confidence == pytest.approx(want, abs=1e-15)
    # anchors of the form: at, 6 dB over, far over
    # a track is born with its candidate's value and recomputes it from its
    # own (here unsmoothed) SNR at every hit after
    for snr, want in ((thr - 0.1, 0.0), (thr, 0.5), (thr + 6.0, 1.0 - 0.5 / math.e),
                      (thr + 60.0, 1.0 - 0.5 * math.exp(-10.0))):
        t2, _ = _tracker()
        c = _candidate(10, 12, CENTER + 11 * BW, snr)
        c.margin_confidence = 0.123
        assert _step(t2, [c], 0)[0].margin_confidence == 0.123
        tr = _step(t2, [c], 1)
        assert tr[0].margin_confidence == pytest.approx(want, abs=1e-15)


def test_band_levels_follow_the_frames_and_restart_when_the_band_moves():
    t, _ = _tracker()
    for k in range(3):  # born on the third hit
        tracks = _step(t, [_candidate(500, 509, CENTER + 504.5 * BW)], k)
    assert tracks[0].state == 1  # live
    g = np.random.default_rng(9)
    levels = []
    for f in range(25):
        frame = (-60.0 + g.normal(0.0, 3.0, BINS)).astype(np.float32)
        t.measure_levels(frame)
        p = np.exp2(frame[500:510].astype(np.float32) * np.float32(0.332192809488736234787))
        levels.append(10.0 * math.log10(float(np.sum(p.astype(np.float64)))))
    tr = t.tracks()[0]
    assert tr.level_frames == 25
    # within float exp2's last place (numpy's and the C library's differ by an
    # ulp now and then): a wrong scale or sum is off by decibels
    assert tr.level_mean_db == pytest.approx(np.mean(levels), abs=1e-6)
    assert math.sqrt(tr.level_m2 / (tr.level_frames - 1)) == pytest.approx(np.std(levels, ddof=1), abs=1e-6)
    # the band moves: the statistics start again over the new bins
    _step(t, [_candidate(501, 511, CENTER + 506 * BW)], 3)
    tr = t.tracks()[0]
    assert (tr.level_first_bin, tr.level_last_bin, tr.level_frames) == (501, 511, 0)
    # a pending track is not measured
    t2, _ = _tracker()
    _step(t2, [_candidate(500, 509, CENTER + 504.5 * BW)], 0)
    t2.measure_levels(np.zeros(BINS, dtype=np.float32))
    assert t2.tracks()[0].level_frames == 0


def test_grid_channel_moves_only_past_the_hysteresis():
    t, cfg = _tracker(grid_channels=M)
    assert cfg.channel_hysteresis == pytest.approx(0.15)
    width = RATE / M
    seen = []
    # the centre walks up across the 10 | 11 boundary and back
    for k, x in enumerate((10.3, 10.6, 10.64, 10.7, 10.4, 10.36, 10.3)):
        tr = _step(t, [_candidate(100, 102, CENTER + x * width)], k)[0]
        seen.append((tr.channel_index, tr.channel, tr.channel_valid))
    assert [s[0] for s in seen] == [10, 10, 10, 11, 11, 11, 10]
    assert all(v == 1 for _, _, v in seen)
    # below the centre: the index is negative and the channel wraps into [0, M)
    t2, _ = _tracker(grid_channels=M)
    tr = _step(t2, [_candidate(100, 102, CENTER - 3.2 * width)], 0)[0]
    assert (tr.channel_index, tr.channel) == (-3, M - 3)
    # no grid: no channel
    t3, _ = _tracker()
    tr = _step(t3, [_candidate(100, 102, CENTER + 7 * width)], 0)[0]
    assert (tr.channel_valid, tr.channel, tr.channel_index) == (0, 0, 0)


def _shapes(excess, bands):
    """Candidates of a frame whose floor is 1 and whose average is 1 + excess."""
    cfg = events.event_config(BW, ZERO, RATE, CENTER)
    floor = np.ones(BINS)
    return events.candidates(cfg, floor + excess, floor, [(s, w, 0.0) for s, w in bands])


def test_band_shapes_by_their_definitions():
    """measure_band (core/detect/shape.h:275-372): peak over mean, the best
    three adjacent bins' share, the lower half's share of the halves, and the
    skirt bins over a tenth of the mean within twice the width and the band's
    own region, on bands whose answers are fixed by construction."""
    x = np.zeros(BINS)
    x[1000:1040] = 2.0                          # flat, 40 bins
    x[2000:2003] = [1.0, 10.0, 1.0]             # a carrier over three bins
    x[3000:3020] = np.arange(1.0, 21.0)         # a ramp ...
    x[3020:3025] = 2.0                          # ... with a 5-bin skirt past its band
    flat, carrier, ramp = _shapes(x, [(1000, 40), (1996, 10), (3000, 20)])
    # flat: the occupied band is all 40 bins
    assert (flat.first_bin, flat.last_bin, flat.shape_measured) == (1000, 1039, 1)
    assert flat.shape_peak_to_mean == pytest.approx(1.0, abs=1e-12)
    assert flat.shape_concentration == pytest.approx(6.0 / 80.0, abs=1e-12)
    assert flat.shape_lower_fraction == pytest.approx(0.5, abs=1e-12)
    assert flat.shape_skirt_fraction == 0.0
    assert flat.shape_skirt_bins_available == 80 + 80   # twice the width each side, room enough
    # carrier: three bins or fewer are wholly concentrated
    assert (carrier.first_bin, carrier.last_bin) == (2000, 2002)
    assert carrier.shape_peak_to_mean == pytest.approx(10.0 / 4.0, abs=1e-12)
    assert carrier.shape_concentration == 1.0
    assert carrier.shape_lower_fraction == pytest.approx(0.5, abs=1e-12)
    # ramp: the occupied band drops its first bin (1 of 210 is under the
    # half-percent tail); 19 bins of 2..20, mean 11, halves of 9
    assert (ramp.first_bin, ramp.last_bin) == (3001, 3019)
    assert ramp.shape_peak_to_mean == pytest.approx(20.0 / 11.0, abs=1e-12)
    assert ramp.shape_concentration == pytest.approx((18 + 19 + 20) / 209.0, abs=1e-12)
    assert ramp.shape_lower_fraction == pytest.approx(54.0 / (54.0 + 144.0), abs=1e-12)
    # below: bin 3000 holds 1, under 1.1; above: the five skirt bins hold 2
    assert ramp.shape_skirt_fraction == pytest.approx(5.0 / 19.0, abs=1e-12)


def test_a_skirt_walk_stops_at_the_neighbouring_band():
    x = np.zeros(BINS)
    x[3400:3410] = 4.0
    x[3410:3420] = 4.0 * 0.5     # a skirt ten bins long ...
    x[3412:3416] = 8.0           # ... with another band starting two bins past the first
    a, b = _shapes(x, [(3400, 10), (3412, 4)])[:2]
    assert (a.first_bin, a.last_bin) == (3400, 3409)
    # its own region ends where the next band starts: two bins of room above
    assert a.shape_skirt_bins_available == 20 + 2
    assert a.shape_skirt_fraction == pytest.approx(2.0 / 10.0, abs=1e-12)
    # a track takes its candidate's shape as measured
    t, _ = _tracker()
    tr = _step(t, [a], 0)[0]
    assert (tr.shape_peak_to_mean, tr.shape_skirt_fraction, tr.shape_measured) == \
        (a.shape_peak_to_mean, a.shape_skirt_fraction, 1)
