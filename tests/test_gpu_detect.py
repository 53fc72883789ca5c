"""GPU vs oracle, bit for bit: the plain detector's decision.

core/detect/detector.cpp estimate_noise_floor (:794) and find_candidates
(:879) on averaged linear spectra: the per-bin floor and spread, the number of
ladder peaks, and the accepted bands (start, width, deflection) must equal the
oracle's exactly.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, rng, to_dev
from oracle import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

BIN_WIDTH = 122.88e6 / 4096 / 16


def _spectrum(kind, bins, seed):
    g = rng(seed)
    if kind == "flat":
        return np.full(bins, 0.25)
    if kind == "raw":
        avg = g.exponential(1.0, bins) * 3e-9
    else:
        avg = g.gamma(64.0, 1.0 / 64.0, bins) * 1e-6
    for _ in range(max(1, bins // 4000)):
        w = int(g.integers(1, max(2, bins // 30)))
        s = int(g.integers(0, bins - w))
        avg[s:s + w] += avg[s:s + w].mean() * float(g.uniform(0.05, 30.0))
    return avg


CASES = [("averaged", 65536, 1), ("averaged", 65536, 2), ("raw", 65536, 3), ("flat", 65536, 4),
         ("averaged", 1000, 5), ("averaged", 4096, 6), ("raw", 777, 7)]


@pytest.mark.parametrize("kind,bins,seed", CASES, ids=[f"{k}-{b}-{s}" for k, b, s in CASES])
@pytest.mark.parametrize("threshold_db", [6.0, 0.0])
def test_detect_matches_oracle(dev, kind, bins, seed, threshold_db):
    from revenant_amd.detect import Detector, config

    avg = _spectrum(kind, bins, 0xDE7EC70000000100 + seed)
    cfg = config(bins, BIN_WIDTH, threshold_db)
    det = Detector(dev, cfg)
    bands, peaks = det.decide(to_dev(avg))
    floor, sigma = det.floor()
    ob, of, osg, opk = oracle.detect(oracle.detect_config(bins, BIN_WIDTH, threshold_db), avg)
    assert_bit_exact(floor, of, what=f"floor {kind} {bins}")
    assert_bit_exact(sigma, osg, what=f"sigma {kind} {bins}")
    assert peaks == opk
    assert len(bands) == len(ob), (bands[:5], ob[:5])
    for (s, w, d), (os_, ow, od) in zip(bands, ob):
        assert (s, w) == (os_, ow) and np.float64(d).tobytes() == np.float64(od).tobytes()
    det.close()


def test_detect_small_budget_and_knots(dev):
    """A small peak budget keeps the strongest peaks (the reference's best-N
    set); few knots and a narrow ladder exercise the other branches."""
    from revenant_amd.detect import Detector, config

    bins = 20000
    avg = _spectrum("averaged", bins, 0xDE7EC70000000200)
    kw = dict(max_peaks=7, noise_knots=3, noise_window_knots=2, max_sum_bins=64,
              split_gap_bins=3, edge_floor_fraction=0.5)
    det = Detector(dev, config(bins, BIN_WIDTH, 3.0, **kw))
    bands, peaks = det.decide(to_dev(avg))
    ob, _, _, opk = oracle.detect(oracle.detect_config(bins, BIN_WIDTH, 3.0, **kw), avg)
    assert peaks == opk and peaks > 7
    assert [(s, w) for s, w, _ in bands] == [(s, w) for s, w, _ in ob]
    det.close()


def test_detect_allow_mask(dev):
    """A mask that allows every bin changes nothing; one that allows only the
    signals' neighbourhoods still finds them, and nothing far from them."""
    from revenant_amd.detect import Detector, config

    bins = 65536
    g = rng(0xDE7EC70000000300)
    avg = g.gamma(64.0, 1.0 / 64.0, bins) * 1e-6
    boxes = [(3000, 50), (30000, 700), (60000, 5)]
    for s, w in boxes:
        avg[s:s + w] += 4e-6
    det = Detector(dev, config(bins, BIN_WIDTH))
    full, _ = det.decide(to_dev(avg))
    ones = torch.arange(bins + 1, dtype=torch.int32, device="cuda")
    same, _ = det.decide(to_dev(avg), ones)
    assert same == full
    allow = np.zeros(bins, dtype=np.int32)
    for s, w in boxes:
        allow[max(0, s - 64):s + w + 64] = 1
    prefix = np.concatenate([[0], np.cumsum(allow)]).astype(np.int32)
    masked, _ = det.decide(to_dev(avg), to_dev(prefix))
    for s, w in boxes:
        assert any(b <= s + w // 2 < b + bw for b, bw, _ in masked), (s, w, masked)
    for b, bw, _ in masked:
        assert allow[max(0, b - bw):b + 2 * bw].any()
    det.close()


def test_detect_on_chain_spectrum(dev):
    """The averaged spectrum of the fused chain (a 4096-channel grid over a
    carrier scene) is decided identically by the GPU and the oracle."""
    from revenant_amd import synth
    from revenant_amd.detect import Detector, config
    from revenant_amd.dsp import FORMAT_CS16
    from revenant_amd.engine import Engine
    from revenant_amd.dsp import GridParams, design_prototype

    rate = 122_880_000
    grid = GridParams(4096, 16, 2048)
    n = 1 << 22
    packed = torch.empty(4 * n, dtype=torch.uint8, device="cuda")
    synth.render_device(dev, synth.wideband_scene(rate, 60, 77), rate, 0, n, noise_dbfs=-60.0,
                        seed=77, fmt=FORMAT_CS16, out=packed)
    eng = Engine(dev, grid, design_prototype(grid, beta=8.0).taps, FORMAT_CS16, max_push=n,
                 ring_blocks=4096, frame_blocks=4, row_frames=4)
    eng.push(packed)
    spec = eng.spectrum().astype(np.float64)
    eng.close()
    det = Detector(dev, config(4096, rate / 4096))
    bands, peaks = det.decide(to_dev(spec))
    ob, _, _, opk = oracle.detect(oracle.detect_config(4096, rate / 4096), spec)
    assert peaks == opk and [(s, w) for s, w, _ in bands] == [(s, w) for s, w, _ in ob]
    assert len(bands) >= 5
    det.close()


def _tie_scenes():
    bins = 4096
    plateau = np.ones(bins)
    plateau[1000:1400] += 4.0
    boxes = np.ones(bins)
    for k in range(12):
        boxes[500 + 100 * k:516 + 100 * k] += 4.0
    return [("plateau", plateau, dict(max_sum_bins=64)),
            ("equal-deflections", boxes, dict(max_sum_bins=64, max_peaks=5)),
            ("equal-deflections-mirrored", boxes[::-1].copy(), dict(max_sum_bins=64,
                                                                     max_peaks=5))]


@pytest.mark.parametrize("name", ["plateau", "equal-deflections", "equal-deflections-mirrored"])
def test_detect_tie_rules_match_oracle(dev, name):
    """The closed-form scenes that pin the twin's plateau middle and its order
    at equal deflection (tests/test_oracle_detect.py), on the device: same
    peaks, same bands, same deflections."""
    from revenant_amd.detect import Detector, config

    (_, avg, kw), = [s for s in _tie_scenes() if s[0] == name]
    det = Detector(dev, config(avg.size, BIN_WIDTH, 0.0, **kw))
    bands, peaks = det.decide(to_dev(avg))
    ob, _, _, opk = oracle.detect(oracle.detect_config(avg.size, BIN_WIDTH, 0.0, **kw), avg)
    assert peaks == opk
    assert [(s, w) for s, w, _ in bands] == [(s, w) for s, w, _ in ob]
    assert [np.float64(d).tobytes() for _, _, d in bands] == \
        [np.float64(d).tobytes() for _, _, d in ob]
    det.close()
