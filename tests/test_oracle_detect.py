"""Pins for the plain detector's decision twin (oracle/detect.cpp).

core/detect/detector.cpp estimate_noise_floor (:794) and find_candidates
(:879), checked against what fixes them independently: a brute-force sort
and mean for each knot's floor, the closed-form deflection of a box on a flat
floor, and boxes of known width coming back as bands on noise.
"""
import math

import numpy as np
import pytest

from helpers import rng
from oracle import oracle

BINS = 65536
BIN_WIDTH = 122.88e6 / 4096 / 16  # a 4096-channel grid at 16 bins per channel


def _knot_floor_brute(avg, k, knots=32, window_knots=8, p_lo=0.10, p_hi=0.40, sigma=4.0):
    bins = avg.size
    spacing = max(1, bins // knots)
    window = min(bins, spacing * window_knots)
    half = window // 2
    centre = int((2 * k + 1) * bins / (2 * knots))
    lo, hi = max(0, centre - half) if centre > half else 0, min(bins, centre + half + 1)
    w = np.sort(avg[lo:hi])
    n = w.size
    r_lo = min(int(p_lo * n), n - 1)
    r_hi = min(max(int(p_hi * n), r_lo), n - 1)
    from statistics import NormalDist
    q_lo, q_hi = NormalDist().inv_cdf(p_lo), NormalDist().inv_cdf(p_hi)
    dev = max(0.0, (w[r_hi] - w[r_lo]) / (q_hi - q_lo))
    level = w[r_hi] - q_hi * dev
    ceiling = level + sigma * dev
    keep = avg[lo:hi][avg[lo:hi] <= ceiling]
    return (keep.mean() if keep.size else level), dev, centre


def test_knot_floors_match_brute_force():
    # an averaged noise floor (64 looks) near 1e-6
    avg = rng(0xDE7EC7000001).gamma(64.0, 1.0 / 64.0, BINS) * 1e-6
    avg[30000:30400] += 5e-5  # a signal the excision must leave out
    cfg = oracle.detect_config(BINS, BIN_WIDTH)
    _, floor, sigma, _ = oracle.detect(cfg, avg)
    for k in range(32):
        f, dev, centre = _knot_floor_brute(avg, k)
        # knot positions are whole bins here, so the interpolated floor
        # equals the knot's own value there
        assert floor[centre] == pytest.approx(f, rel=1e-12), k
        # Acklam's quantile approximation vs the exact inverse CDF: 1e-9
        assert sigma[centre] == pytest.approx(dev, rel=1e-7), k
    # the noise mean is 1e-6; the excised floor sits near it
    assert np.all(np.abs(floor / 1e-6 - 1.0) < 0.05)


@pytest.mark.parametrize("width,start", [(64, 1000), (256, 40000), (1024, 7)])
def test_box_on_flat_floor_closed_form(width, start):
    """Floor 1 everywhere, excess E on [start, start+width): the knot floors
    are exactly 1 (every window's percentiles are 1, so the spread is 0 and
    the box is excised), the best rung is the box's own width with deflection
    E*sqrt(W), and growth at level 0 stops exactly at the box's edges."""
    E = 0.75
    avg = np.ones(BINS)
    avg[start:start + width] += E
    cfg = oracle.detect_config(BINS, BIN_WIDTH, threshold_db=0.0)
    bands, floor, sigma, _ = oracle.detect(cfg, avg)
    assert np.all(floor == 1.0) and np.all(sigma == 0.0)
    assert len(bands) == 1
    s, w, d = bands[0]
    assert (s, w) == (start, width)
    assert d == float(np.float32(E * width * math.sqrt(width) / width))


def test_flat_spectrum_has_no_bands():
    cfg = oracle.detect_config(BINS, BIN_WIDTH)
    # a power of two, so every window's mean is exact
    bands, floor, _, peaks = oracle.detect(cfg, np.full(BINS, 0.25))
    assert bands == [] and peaks == 0 and np.all(floor == 0.25)


def test_boxes_on_noise_come_back_as_bands():
    g = rng(0xDE7EC7000002)
    avg = g.exponential(1.0, BINS) / 64.0 + 1.0  # an averaged noise floor near 1
    boxes = [(2000, 40, 2.0), (9000, 300, 0.5), (20000, 2000, 0.2), (50000, 8, 10.0)]
    for s, w, e in boxes:
        avg[s:s + w] += e
    bands, _, _, _ = oracle.detect(oracle.detect_config(BINS, BIN_WIDTH), avg)
    for s, w, _ in boxes:
        hit = [b for b in bands if b[0] <= s + w // 2 < b[0] + b[1]]
        assert len(hit) == 1, (s, w, bands[:10])
        b0, bw, _ = hit[0]
        overlap = min(b0 + bw, s + w) - max(b0, s)
        assert overlap >= 0.9 * w and bw <= 1.5 * w + 16, (s, w, hit)


def test_a_plateau_seeds_at_its_middle():
    """A flat box wider than the widest rung: every window of that rung wholly
    inside it has the same deflection (exact sums: floor 1, excess 4), so the
    maximum is a plateau and the seed is its middle (detector.cpp:1002-1012).
    Growth then reaches one seed width each side, so the band is centred on
    the box: [1104, 1296) for the box [1000, 1400) at rung 64. A seed taken at
    the plateau's first or last position would give [1000, 1128) or
    [1272, 1400)."""
    bins = 4096
    avg = np.ones(bins)
    avg[1000:1400] += 4.0
    cfg = oracle.detect_config(bins, BIN_WIDTH, threshold_db=0.0, max_sum_bins=64)
    bands, floor, sigma, _ = oracle.detect(cfg, avg)
    assert np.all(floor == 1.0) and np.all(sigma == 0.0)
    assert [(s, w) for s, w, _ in bands] == [(1104, 192)]
    assert bands[0][2] == float(np.float32(4.0 * 64 * 8 / 64))


def test_equal_deflections_keep_the_lowest_starts_when_the_budget_binds():
    """Twelve identical boxes give twelve peaks of exactly equal deflection
    (16 bins of excess 4: 4 * 16 * 4 / 16 = 16); with a budget of five, the
    order that breaks the tie (strongest, then lower start, then narrower;
    detector.cpp:951-960) keeps the five lowest starts, whichever order the
    scan met them in, and the accepted bands are those boxes."""
    bins = 4096
    avg = np.ones(bins)
    starts = [500 + 100 * k for k in range(12)]
    for s in starts:
        avg[s:s + 16] += 4.0
    cfg = oracle.detect_config(bins, BIN_WIDTH, threshold_db=0.0, max_sum_bins=64, max_peaks=5)
    bands, _, _, peaks = oracle.detect(cfg, avg)
    assert [(s, w) for s, w, _ in bands] == [(s, 16) for s in starts[:5]]
    assert all(d == 16.0 for _, _, d in bands)
    # mirrored, the same rule keeps the five lowest starts of the mirror image
    bands_m, _, _, _ = oracle.detect(cfg, avg[::-1].copy())
    mirrored = sorted(bins - s - 16 for s in starts)
    assert [(s, w) for s, w, _ in bands_m] == [(s, 16) for s in mirrored[:5]]
