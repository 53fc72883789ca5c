"""Pins for the oracle's detection-event stages (oracle/detect.cpp): candidates
from accepted bands, residual rejection and the tracker
(core/detect/detector.cpp:1168-1900), each against what follows from the
rules by hand: a box's bins and hertz, the occupied-band trim, the split at a
floor-level gap, a burst's birth, hold and drop at known decisions (its
event boundaries in sample indices), a merge and the split after it, and an
exponentially emptying band withheld after the configured run.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle

BINS = 4096
BW = 1000.0              # Hz per bin
ZERO = -BINS / 2 * BW    # bin 0's centre
RATE = 1_000_000         # samples per second
STEP = 100_000           # samples between decisions (0.1 s)


def cfg(**kw):
    return oracle.event_config(BW, ZERO, RATE, **kw)


def box(spec):
    """floor 1, excess per (start, width, level)."""
    avg = np.ones(BINS)
    for s, w, e in spec:
        avg[s:s + w] += e
    return avg, np.ones(BINS)


def test_a_box_is_one_candidate_with_its_edges_in_hertz():
    avg, flo = box([(1000, 40, 3.0)])
    (c,) = oracle.candidates(cfg(), avg, flo, [(996, 48, 0.0)])
    assert (c.first_bin, c.last_bin) == (1000, 1039)
    low, high = ZERO + (1000 - 0.5) * BW, ZERO + (1039 + 0.5) * BW
    assert c.center == round(0.5 * (low + high)) and c.bandwidth == round(high - low) == 40_000
    assert c.snr_2500_db == 10 * math.log10(3.0 * 40 * BW / (1.0 * 2500.0))
    assert c.noise_floor_dbfs == 0.0 and c.peak_dbfs == 10 * math.log10(4.0)


def test_the_occupied_band_drops_tails_under_half_a_percent_each():
    # shoulders of 1e-4 per bin: 3 bins carry 3e-4 < tail = 20.0006 * 0.005
    avg, flo = box([(2000, 3, 1e-4), (2003, 20, 1.0), (2023, 3, 1e-4)])
    (c,) = oracle.candidates(cfg(), avg, flo, [(2000, 26, 0.0)])
    assert (c.first_bin, c.last_bin) == (2003, 2022)


@pytest.mark.parametrize("gap,parts", [(8, 2), (12, 2), (7, 1), (3, 1)])
def test_a_floor_level_gap_of_split_gap_bins_cuts_the_band(gap, parts):
    avg, flo = box([(3000, 30, 2.0), (3030 + gap, 30, 2.0)])
    got = oracle.candidates(cfg(), avg, flo, [(3000, 60 + gap, 0.0)])
    assert len(got) == parts
    if parts == 2:
        assert [(c.first_bin, c.last_bin) for c in got] == [(3000, 3029),
                                                           (3030 + gap, 3059 + gap)]


def test_a_band_under_the_threshold_is_no_candidate():
    avg, flo = box([(500, 2, 0.5)])  # 0.5 * 2 * 1000 / 2500 = 0.4 < 6 dB
    assert oracle.candidates(cfg(), avg, flo, [(500, 2, 0.0)]) == []


def _run(tracker, frames, start=0):
    """frames: per decision, the box list; returns the tracks after each."""
    out = []
    for k, spec in enumerate(frames):
        avg, flo = box(spec)
        bands = [(s, w, 0.0) for s, w, _ in spec]
        cands = oracle.candidates(tracker_cfg, avg, flo, bands)
        tracker.update(avg, flo, cands, start + k * STEP, 0.1 if k or start else 0.0)
        out.append(tracker.tracks())
    return out


tracker_cfg = cfg(residual_decisions=0)


def test_a_burst_is_born_held_and_dropped_at_the_decisions_the_rules_give():
    t = oracle.OracleTracker(tracker_cfg, BINS)
    frames = [[(1000, 40, 3.0)]] * 10 + [[]] * 35
    hist = _run(t, frames)
    assert [tr.state for tr in hist[0]] == [0] and hist[0][0].hits == 1  # pending
    assert hist[1][0].state == 0 and hist[2][0].state == 1               # born at the 3rd hit
    tr = hist[9][0]
    assert (tr.id, tr.hits, tr.first_seen, tr.last_detected) == (1, 10, 0, 9 * STEP)
    assert tr.confidence == pytest.approx(1 - 0.65 ** 10, rel=1e-12)
    # held with confidence halving every second, until the 3 s hold runs out:
    # silent 31 decisions = 3.1 s > 3 s at decision 40
    held = hist[39][0]
    assert held.state == 2 and held.misses == 30 and held.last_detected == 9 * STEP
    assert held.last_seen == 39 * STEP and held.held_spells == 1
    assert held.confidence == pytest.approx((1 - 0.65 ** 10) * 2 ** (-0.1 * 30), rel=1e-9)
    assert hist[40] == []
    t.close()


def test_a_pending_track_that_misses_once_is_discarded():
    t = oracle.OracleTracker(tracker_cfg, BINS)
    hist = _run(t, [[(1000, 40, 3.0)]] * 2 + [[]] + [[(1000, 40, 3.0)]])
    assert hist[1][0].hits == 2 and hist[2] == []
    assert hist[3][0].id == 2 and hist[3][0].first_seen == 3 * STEP  # a new event
    t.close()


def test_two_tracks_merge_into_the_elder_and_split_again():
    t = oracle.OracleTracker(tracker_cfg, BINS)
    a, b = (1000, 20, 3.0), (1030, 20, 3.0)
    wide = (1000, 50, 3.0)
    frames = [[a]] * 3 + [[a, b]] * 3 + [[wide]] * 2 + [[a, b]] * 2
    hist = _run(t, frames)
    ids = {tr.id: tr for tr in hist[5]}
    assert set(ids) == {1, 2} and all(tr.state == 1 for tr in ids.values())
    merged = {tr.id: tr for tr in hist[6]}
    # the wide candidate goes to the elder (id 1, first seen at decision 0);
    # the younger becomes its child
    assert merged[1].state == 1 and merged[2].state == 3 and merged[2].merged_into == 1
    assert merged[2].last_detected == 5 * STEP and merged[2].last_seen == 6 * STEP
    split = {tr.id: tr for tr in hist[8]}
    assert split[1].state == 1 and split[2].state == 1 and split[2].merged_into == 0
    assert split[2].first_seen == 3 * STEP  # the same event, not a new one
    t.close()


def test_an_emptying_average_is_withheld_after_the_run_and_a_steady_band_is_not():
    """A band whose mean excess falls exactly as a 1 s average empties
    (10/ln 10 dB per second) is withheld once residual_decisions = 4 falling
    decisions in a row have been seen; a steady band beside it never is."""
    c = cfg(residual_decisions=4)
    t = oracle.OracleTracker(c, BINS)
    kept = []
    for k in range(12):
        e = 50.0 if k < 4 else 50.0 * math.exp(-0.1 * (k - 3))
        avg, flo = box([(1000, 64, e), (3000, 64, 50.0)])
        cands = oracle.candidates(c, avg, flo, [(1000, 64, 0.0), (3000, 64, 0.0)])
        assert len(cands) == 2
        kept.append(t.update(avg, flo, cands, k * STEP, 0.1))
    # decisions 4..6 fall (runs 1..3) and are kept; from the 4th fall on the
    # decaying band is withheld, the steady one always kept
    assert kept == [2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1]
    t.close()
