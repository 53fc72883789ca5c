"""The product's detection-event stages (rvx_candidates, rvx_tracker_*, host
code behind the C ABI) against the oracle's, bit for bit: candidates of every
field, the residual rejection's kept count and every track's integers and
doubles, over seeded sequences of decisions where signals appear, drift,
merge, split, fade exponentially and vanish, on the oracle's decisions
(the GPU decision is the same bands and floor; tests/test_gpu_detect.py)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import rng
from oracle import oracle
from revenant_amd import events

BINS = 8192
BW = 122.88e6 / 4096 / 16
ZERO = -BINS / 2 * BW
RATE = 122_880_000
STEP = RATE // 10


def _frames(seed, decisions=60):
    g = rng(seed)
    emitters = []
    for _ in range(14):
        s = int(g.integers(50, BINS - 400))
        w = int(g.integers(4, 300))
        e = float(np.exp(g.uniform(np.log(0.05), np.log(20.0))))
        on = int(g.integers(0, decisions // 2))
        off = on + int(g.integers(3, decisions))
        drift = int(g.integers(-2, 3))
        emitters.append((s, w, e, on, off, drift))
    for k in range(decisions):
        avg = g.gamma(64.0, 1.0 / 64.0, BINS)  # an averaged noise floor near 1
        for s, w, e, on, off, drift in emitters:
            s2 = s + drift * k
            if on <= k < off:
                level = e
            elif k >= off:
                level = e * np.exp(-0.1 * (k - off + 1))  # the average emptying
            else:
                continue
            avg[s2:s2 + w] += level
        yield avg


def _same_candidates(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert (x.first_bin, x.last_bin, x.center_hz, x.bandwidth_hz) == \
            (y.first_bin, y.last_bin, y.center, y.bandwidth)
        for f in ("snr_2500_db", "noise_floor_dbfs", "peak_dbfs"):
            assert np.float64(getattr(x, f)).tobytes() == np.float64(getattr(y, f)).tobytes(), f


def _same_tracks(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        ints = ("id", "state", "hits", "misses", "held_spells", "first_seen", "last_seen",
                "last_detected", "merged_into", "first_bin", "last_bin")
        assert tuple(getattr(x, f) for f in ints) == tuple(getattr(y, f) for f in ints)
        assert (x.center_hz, x.bandwidth_hz) == (y.center, y.bandwidth)
        for f in ("snr_2500_db", "confidence"):
            assert np.float64(getattr(x, f)).tobytes() == np.float64(getattr(y, f)).tobytes(), f


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("residual", [4, 0])
def test_events_match_the_oracle_decision_by_decision(seed, residual):
    dcfg = oracle.detect_config(BINS, BW)
    ocfg = oracle.event_config(BW, ZERO, RATE, 96_000_000, residual_decisions=residual)
    pcfg = events.event_config(BW, ZERO, RATE, 96_000_000, residual_decisions=residual)
    ot, pt = oracle.OracleTracker(ocfg, BINS), events.Tracker(pcfg, BINS)
    saw_tracks = saw_merge = 0
    for k, avg in enumerate(_frames(0xE7E47500 + seed)):
        bands, floor, _, _ = oracle.detect(dcfg, avg)
        oc = oracle.candidates(ocfg, avg, floor, bands)
        pc = events.candidates(pcfg, avg, floor, bands)
        _same_candidates(pc, oc)
        now = 5_000_000 + k * STEP
        assert pt.update(avg, floor, pc, now, 0.1) == ot.update(avg, floor, oc, now, 0.1)
        tr = pt.tracks()
        _same_tracks(tr, ot.tracks())
        saw_tracks = max(saw_tracks, sum(t.state == 1 for t in tr))
        saw_merge |= any(t.state == 3 for t in tr)
    assert saw_tracks >= 3
    st = pt.stats()
    assert st.decisions == 60 and st.tracks_born >= 3
    if residual:
        assert st.candidates_residual > 0
    ot.close()
    pt.close()


def test_refusals():
    pcfg = events.event_config(BW, ZERO, RATE)
    avg = np.ones(BINS)
    with pytest.raises(Exception):
        events.candidates(pcfg, avg, avg, [(BINS - 4, 8, 0.0)])  # past the last bin
    with pytest.raises(Exception):
        events.candidates(pcfg, avg, avg, [(100, 8, 0.0), (104, 8, 0.0)])  # overlapping
    bad = events.event_config(BW, ZERO, RATE, birth_hits=0)
    with pytest.raises(Exception):
        events.Tracker(bad, BINS)
