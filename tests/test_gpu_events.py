"""The detection events on the GPU decision: the averaged spectrum goes
through the GPU detector (floor and bands, core/detect/detector.cpp:794/879)
and then through the product's candidate and tracker stages, decision after
decision; the oracle runs the same sequence on its own decision. Candidates
and tracks must agree bit for bit."""
import numpy as np
import pytest

from helpers import to_dev
from oracle import oracle
from revenant_amd import events

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import test_events_parity as tp  # noqa: E402


@pytest.mark.parametrize("seed", [5, 6])
def test_gpu_decisions_drive_the_same_events(dev, seed):
    from revenant_amd.detect import Detector, config

    det = Detector(dev, config(tp.BINS, tp.BW))
    dcfg = oracle.detect_config(tp.BINS, tp.BW)
    ocfg = oracle.event_config(tp.BW, tp.ZERO, tp.RATE, 96_000_000)
    pcfg = events.event_config(tp.BW, tp.ZERO, tp.RATE, 96_000_000)
    ot, pt = oracle.OracleTracker(ocfg, tp.BINS), events.Tracker(pcfg, tp.BINS)
    born = 0
    for k, avg in enumerate(tp._frames(0xE7E47500 + seed)):
        bands, _ = det.decide(to_dev(avg))
        floor, _ = det.floor()
        ob, of, _, _ = oracle.detect(dcfg, avg)
        oc = oracle.candidates(ocfg, avg, of, ob)
        pc = events.candidates(pcfg, avg, floor, bands)
        tp._same_candidates(pc, oc)
        now = 5_000_000 + k * tp.STEP
        assert pt.update(avg, floor, pc, now, 0.1) == ot.update(avg, of, oc, now, 0.1)
        tp._same_tracks(pt.tracks(), ot.tracks())
    born = pt.stats().tracks_born
    assert born >= 3
    det.close()
    ot.close()
    pt.close()
