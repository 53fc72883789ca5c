"""Pins the front-end correction's oracle and the host estimator.

Oracle: the chunk moments against float64 sums (exactly, on integer-valued
samples whose every partial sum is exact in float), the correction undoing the
imbalance model to 1e-6 (tests/reference/test_iq_correction.cpp:166-199) and
the identity leaving samples alone. Estimator (the product's host code, fed by
the oracle's per-sample work on the reference's schedule, blocks of 32768 and
three frames in flight): tests/engine/test_front_end_correction.cpp:179-260 --
a -33 dBFS DC spike down by 40 dB with a tone 5 kHz away unchanged, image
rejection from 28.9 dB to past 60 dB with the gain and phase read back, both
at once, and an implausible imbalance measured but not applied.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.front_end import FrontEndCorrector, total_moments

RATE = 2_400_000
TONE_HZ = 5000
MEASURE = 480 * 2500          # a whole number of 5 kHz cycles, half a second
STREAM = 2 * RATE
BLOCK = 32768
IN_FLIGHT = 3
DC = (0.02, 0.01)             # -33.0 dBFS
GAIN, PHASE = 1.05, 3.0 * math.pi / 180.0


def test_moments_match_float64_sums_and_are_exact_on_integers():
    g = np.random.default_rng(11)
    cap = 1 << 12
    ring = g.standard_normal(2 * cap).astype(np.float32)
    off, count = cap - 300, 2000  # wraps, ends in a partial chunk
    got = oracle.iq_moments(cap - 1, off, count, ring)
    z = ring.reshape(-1, 2)[(off + np.arange(count)) % cap].astype(np.float64)
    for c in range(got.shape[0]):
        seg = z[c * 64:(c + 1) * 64]
        want = [seg[:, 0].sum(), seg[:, 1].sum(), (seg[:, 0] ** 2).sum(), (seg[:, 1] ** 2).sum(),
                (seg[:, 0] * seg[:, 1]).sum()]
        np.testing.assert_allclose(got[c], want, rtol=1e-5, atol=1e-4)
    ints = g.integers(-100, 101, size=2 * cap).astype(np.float32)
    got = oracle.iq_moments(cap - 1, 5, 1000, ints)
    zi = ints.reshape(-1, 2)[5:1005].astype(np.int64)
    for c in range(got.shape[0]):
        seg = zi[c * 64:(c + 1) * 64]
        want = [seg[:, 0].sum(), seg[:, 1].sum(), (seg[:, 0] ** 2).sum(), (seg[:, 1] ** 2).sum(),
                (seg[:, 0] * seg[:, 1]).sum()]
        assert np.array_equal(got[c], np.array(want, dtype=np.float32)), c
    with pytest.raises(ValueError):
        oracle.iq_moments(cap - 1, 0, cap + 1, ring)


def test_correction_undoes_the_imbalance_model():
    n = np.arange(256)
    ideal = np.exp(1j * 0.1 * n)
    g, phi = 1.05, 3.0 * math.pi / 180.0
    ring = np.stack([ideal.real, g * (ideal.imag * math.cos(phi) + ideal.real * math.sin(phi))],
                    axis=1).astype(np.float32).ravel()
    out = oracle.iq_correct(255, 0, 256, 0.0, 0.0, float(np.float32(-math.tan(phi))),
                            float(np.float32(1.0 / (g * math.cos(phi)))), ring)
    z = out.reshape(-1, 2)
    err = np.abs((z[:, 0] + 1j * z[:, 1]) - ideal.astype(np.complex64))
    assert err.max() < 1e-6
    # the identity leaves every (non-negative-zero) sample's bits alone
    ring2 = np.random.default_rng(3).standard_normal(512).astype(np.float32)
    same = oracle.iq_correct(255, 17, 200, 0.0, 0.0, 0.0, 1.0, ring2)
    assert np.array_equal(same.view(np.uint32), ring2.view(np.uint32))


def render(dc, gain, phase, seed=1):
    """tests/engine/impaired_front_end.h: I_r = I + dc_i + n, Q_r = g (Q cos
    phi + I sin phi) + dc_q + n, a 0.1 tone 5 kHz up, 1e-4 rms noise."""
    t = np.arange(STREAM, dtype=np.float64)
    ang = 2 * math.pi * TONE_HZ * t / RATE
    i, q = 0.1 * np.cos(ang), 0.1 * np.sin(ang)
    r = np.random.default_rng(seed)
    i_r = i + dc[0] + 1e-4 * r.standard_normal(STREAM)
    q_r = gain * (q * math.cos(phase) + i * math.sin(phase)) + dc[1] + \
        1e-4 * r.standard_normal(STREAM)
    return np.stack([i_r, q_r], axis=1).astype(np.float32)


def correct_on_host(stream, dc: bool, iq: bool):
    """The graph's schedule (test_front_end_correction.cpp:104-159): block k's
    moments are read when block k + 3 is recorded; block k is corrected with
    the estimate held before its own moments were taken."""
    cap = 1 << 16
    ring = np.zeros(2 * cap, dtype=np.float32)
    pending = [None] * IN_FLIGHT
    fe = FrontEndCorrector()
    out = np.empty_like(stream)
    off = 0
    for k, first in enumerate(range(0, len(stream), BLOCK)):
        count = min(BLOCK, len(stream) - first)
        slot = k % IN_FLIGHT
        if pending[slot] is not None:
            fe.update_sums(pending[slot][0], pending[slot][1], RATE)
            pending[slot] = None
        idx = (off + np.arange(count)) & (cap - 1)
        ring.reshape(-1, 2)[idx] = stream[first:first + count]
        pending[slot] = (oracle.iq_moments(cap - 1, off, count, ring), count)
        if fe.estimate().measured:
            p = fe.correction(dc, iq)
            ring = oracle.iq_correct(cap - 1, off, count, p.dc_i, p.dc_q, p.cross, p.scale, ring)
        out[first:first + count] = ring.reshape(-1, 2)[idx]
        off = (off + count) & (cap - 1)
    return out, fe.estimate()


def component(stream, f_hz):
    z = stream[-MEASURE:].astype(np.float64)
    n = np.arange(MEASURE) % 480
    w = np.exp(-2j * math.pi * f_hz * n / RATE)
    return np.mean((z[:, 0] + 1j * z[:, 1]) * w)


def db(a):
    return 10 * math.log10(max(abs(a) ** 2, 1e-30))


def test_dc_removal_takes_the_spike_down_40_db_and_leaves_the_tone():
    raw = render(DC, 1.0, 0.0)
    fixed, est = correct_on_host(raw, True, False)
    dc_before, dc_after = db(component(raw, 0.0)), db(component(fixed, 0.0))
    tone_before, tone_after = db(component(raw, TONE_HZ)), db(component(fixed, TONE_HZ))
    assert abs(dc_before - -33.01) < 0.05
    assert dc_before - dc_after >= 40.0, (dc_before, dc_after)
    assert abs(tone_after - tone_before) < 0.001
    assert abs(est.dc_dbfs - -33.01) < 0.05
    off, _ = correct_on_host(raw, False, False)
    assert db(component(off, 0.0)) == dc_before


def test_iq_correction_lifts_image_rejection_past_60_db():
    raw = render((0.0, 0.0), GAIN, PHASE)
    fixed, est = correct_on_host(raw, False, True)
    irr_before = db(component(raw, TONE_HZ)) - db(component(raw, -TONE_HZ))
    irr_after = db(component(fixed, TONE_HZ)) - db(component(fixed, -TONE_HZ))
    assert abs(irr_before - 28.93) < 0.1, irr_before
    assert irr_after >= 60.0, irr_after
    assert est.iq_plausible
    assert abs(est.gain_error_db - 20 * math.log10(GAIN)) < 0.01
    assert abs(est.phase_error_deg - 3.0) < 0.05
    assert abs(est.image_rejection_db - 28.93) < 0.1
    assert abs(db(component(fixed, TONE_HZ)) - -20.0) < 0.01


def test_both_corrections_together():
    raw = render(DC, GAIN, PHASE)
    fixed, _ = correct_on_host(raw, True, True)
    assert db(component(raw, 0.0)) - db(component(fixed, 0.0)) >= 40.0
    assert db(component(fixed, TONE_HZ)) - db(component(fixed, -TONE_HZ)) >= 60.0


def test_an_implausible_imbalance_is_measured_and_left_unapplied():
    fe = FrontEndCorrector()
    m = _lib.IqMomentsTotal(1000, 0.0, 0.0, 500.0, 1.0, 0.0)
    fe.update(m, RATE)
    e = fe.estimate()
    assert e.measured and not e.iq_plausible
    p = fe.correction(True, True)
    assert p.cross == 0.0 and p.scale == 1.0
    # nothing measured yet: the identity, whatever is asked
    fresh = FrontEndCorrector().correction(True, True)
    assert (fresh.dc_i, fresh.dc_q, fresh.cross, fresh.scale) == (0.0, 0.0, 0.0, 1.0)


def test_total_moments_adds_chunks_in_order():
    sums = np.arange(25, dtype=np.float32).reshape(5, 5)
    t = total_moments(sums, 64 * 4 + 1)   # five chunks
    assert t.samples == 257 and t.sum_i == float(sums[:, 0].sum())
    t = total_moments(sums, 64 * 2)       # only two chunks' worth of samples
    assert t.sum_iq == float(sums[:2, 4].sum())
