"""Pins the carrier loop's oracle and its design against the reference's own
measurements (tests/reference/test_vrx_carrier.cpp): the mode table and the
SAM plan (:528-572), the loop gains and the noise bandwidth they deliver
(:574-608), retunes and restarts (:610-662), blocking invariance (:664-682),
lock times (:684-774), tracking a drifting carrier (:776-843) and DSB's unit
level at any carrier phase (:845-901). A slip in the loop (a sign, the
detector, the NCO's direction, a clamp) fails to lock, locks on the mirror
frequency, or misses these bounds.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (CARRIER_COSTAS, CARRIER_DAMPING, CARRIER_NONE, CARRIER_PLL,
                              CARRIER_ACQUIRE_HZ, CARRIER_TRACK_HZ, DEMOD_AM, DEMOD_DSB,
                              DEMOD_SAM, VrxRequest, carrier_check, carrier_loop_gains,
                              carrier_power, carrier_retune, design_carrier_loop, plan_vrx)
from tests.carrier_signals import (GRID, SOURCE_RATE, TONE_HZ, TWO_PI, Signal, am_tone,
                                   audio_between, busy_stream, dsb_plan, dsb_tone, lock_time,
                                   make_plan, nco_for, programme, render, run_chain, sam_plan,
                                   settled_frequency, state_frequency, tone_amplitude)


def test_modes_and_the_sam_plan():
    for mode in range(13):
        want = CARRIER_PLL if mode == DEMOD_SAM else CARRIER_COSTAS if mode == DEMOD_DSB \
            else CARRIER_NONE
        assert carrier_power(mode) == want
    assert carrier_power(DEMOD_SAM + 1) == CARRIER_NONE
    sam, dsb = sam_plan(), dsb_plan()
    am = make_plan(DEMOD_AM, 10_000)
    assert sam.plan.carrier.power == CARRIER_PLL and dsb.plan.carrier.power == CARRIER_COSTAS
    assert am.plan.carrier.power == CARRIER_NONE
    # SAM is AM's receiver but for the detector: channel, DC window, gain
    assert (sam.plan.passband_low, sam.plan.passband_high) == (am.plan.passband_low,
                                                                 am.plan.passband_high)
    assert sam.demod.dc_taps == am.demod.dc_taps and sam.plan.demod_gain == am.plan.demod_gain
    assert sam.demod.mode == DEMOD_SAM and sam.plan.carrier.nco_log2 == sam.fine.nco_log2
    # the integrator's bound is the passband's reach
    for p, reach in ((sam, 5000.0), (dsb, 3000.0)):
        got = float(p.plan.carrier_loop.clamp_radians) * p.plan.demod_rate / TWO_PI
        assert abs(got - reach) < 0.01
    # a coherent detector needs the shared rate floor, AM twice the width
    s = plan_vrx(GRID, SOURCE_RATE, VrxRequest(196_500, DEMOD_SAM, 10_000, audio_rate=5000))
    a = plan_vrx(GRID, SOURCE_RATE, VrxRequest(196_500, DEMOD_AM, 10_000, audio_rate=5000))
    assert s.plan.demod_rate == 15_000 and a.plan.demod_rate == 20_000
    # the plan's loop is design_carrier_loop at its rate and reach, bit for bit
    d = design_carrier_loop(sam.plan.demod_rate, 5000)
    assert bytes(d) == bytes(sam.plan.carrier_loop)


def test_loop_gains_and_the_bandwidth_they_deliver():
    zeta = CARRIER_DAMPING
    k1, k2 = carrier_loop_gains(1.0, zeta, 48_000)
    theta = (1.0 / 48_000.0) / (zeta + 1.0 / (4.0 * zeta))
    assert abs(k1 / (4.0 * zeta * theta) - 1.0) < 1e-4
    assert abs(k2 / (4.0 * theta * theta) - 1.0) < 1e-4
    # noise bandwidth of the linearised loop from its impulse response,
    # B_n = sum(h^2) / (2T); the oscillator updates after the sample it rotated
    for asked in (CARRIER_ACQUIRE_HZ, CARRIER_TRACK_HZ):
        g1, g2 = carrier_loop_gains(asked, zeta, 48_000)
        phase = integ = total = 0.0
        for n in range(480_000):
            err = (1.0 if n == 0 else 0.0) - phase
            integ += g2 * err
            phase += g1 * err + integ
            total += phase * phase
            if n > 2000 and abs(phase) < 1e-30:
                break
        measured = total * 48_000.0 / 2.0
        assert abs(measured / asked - 1.0) < 0.1, (asked, measured)
    assert carrier_loop_gains(0.0, zeta, 48_000) == (0.0, 0.0)
    assert carrier_loop_gains(30.0, zeta, 0) == (0.0, 0.0)
    with pytest.raises(_lib.RvxError):
        design_carrier_loop(0, 5000)
    with pytest.raises(_lib.RvxError):
        design_carrier_loop(48_000, 0)
    p = design_carrier_loop(48_000, 5000)
    assert p.hold_in == 480 and p.hold_out == 9600
    assert np.float32(p.level_alpha) == np.float32(1.0 - math.exp(-1.0 / (0.005 * 48_000)))
    assert np.float32(p.lock_alpha) == np.float32(1.0 - math.exp(-1.0 / (0.020 * 48_000)))
    # a reach past a quarter turn per sample is held at the ceiling
    assert np.float32(design_carrier_loop(8_000, 4_000).clamp_radians) == \
        np.float32(0.5 * math.pi)


def test_nco_table_is_quadrant_symmetric_bit_for_bit():
    """Entry q*N/4 + r is entry r turned by q quarter turns, a zero written as
    +0: the carrier kernel keeps only the first quarter on chip."""
    from revenant_amd.dsp import build_nco_table
    for k in range(2, 21):
        t = build_nco_table(k).view(np.float32).reshape(-1, 2)
        n = 1 << k
        idx = np.arange(n)
        q, r = idx // (n // 4), idx % (n // 4)
        e = t[r]
        x = np.where(q % 2 == 1, e[:, 1], e[:, 0])
        y = np.where(q % 2 == 1, e[:, 0], e[:, 1])
        y = np.where((q == 1) | (q == 2), -y, y)
        x = np.where(q >= 2, -x, x)
        x[x == 0] = 0.0
        y[y == 0] = 0.0
        got = np.stack([x, y], axis=1).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), t.view(np.uint32)), k


def test_retune_rules():
    assert carrier_retune(10_001, 10_000, 48_000)[0]
    assert carrier_retune(-10_001, 10_000, 48_000)[0]
    assert not carrier_retune(10_000, 10_000, 48_000)[0]
    reset, shift = carrier_retune(300, 10_000, 48_000)
    assert not reset and abs(shift + TWO_PI * 300.0 / 48_000.0) < 1e-15


def _valid():
    plan = sam_plan()
    return plan.plan.carrier, plan.carrier_block(1023, 0, 100)


@pytest.mark.parametrize("field,value", [
    ("power", 0), ("power", 3), ("nco_log2", 0), ("nco_log2", 25), ("in_mask", 1000),
    ("count", 1025), ("flags", 2), ("shift", float("inf")), ("clamp_radians", 0.0),
    ("clamp_radians", 1.6), ("k1_acquire", 1.0), ("k1_track", 0.0), ("k2_acquire", 1.0),
    ("k2_track", -0.1), ("level_alpha", 0.0), ("lock_alpha", 1.5), ("hold_in", 0),
    ("hold_out", 0)])
def test_validation_refuses_what_the_reference_refuses(field, value):
    cfg, p = _valid()
    carrier_check(cfg, p)  # the valid block passes
    target = cfg if field in ("power", "nco_log2") else p
    setattr(target, field, value)
    with pytest.raises(_lib.RvxError):
        carrier_check(cfg, p)
    ring = np.zeros(2 * 1024, dtype=np.float32)
    state = np.zeros(8, dtype=np.uint32)
    with pytest.raises(ValueError):
        oracle.vrx_carrier(cfg, p, nco_for(sam_plan()), ring, state)


def test_a_restart_is_a_loop_that_has_seen_nothing():
    plan = sam_plan()
    rate = float(plan.plan.demod_rate)
    a = render(Signal(am_tone(0.3, 0.8), lambda t: np.full_like(t, 140.0)), rate, 9600)
    b = render(Signal(am_tone(0.01, 0.008), lambda t: np.full_like(t, -60.0), 2.2), rate, 9600)
    nco = nco_for(plan)
    mask = (1 << 15) - 1
    ring = np.zeros(2 * (mask + 1), dtype=np.float32)
    ring[:2 * 9600] = a.view(np.float32)
    ring[2 * 9600:2 * 19200] = b.view(np.float32)
    state = np.zeros(8, dtype=np.uint32)
    oracle.vrx_carrier(plan.plan.carrier, plan.carrier_block(mask, 0, 9600), nco, ring, state)
    assert state[5] == 1
    oracle.vrx_carrier(plan.plan.carrier, plan.carrier_block(mask, 9600, 9600, reset=True), nco,
                       ring, state)
    fresh = np.zeros_like(ring)
    fresh[2 * 9600:2 * 19200] = b.view(np.float32)
    fresh_state = np.zeros(8, dtype=np.uint32)
    oracle.vrx_carrier(plan.plan.carrier, plan.carrier_block(mask, 9600, 9600), nco, fresh,
                       fresh_state)
    assert np.array_equal(ring[2 * 9600:2 * 19200].view(np.uint32),
                          fresh[2 * 9600:2 * 19200].view(np.uint32))
    assert np.array_equal(state, fresh_state)


@pytest.mark.parametrize("mode", [DEMOD_SAM, DEMOD_DSB])
def test_output_does_not_depend_on_where_blocks_fall(mode):
    plan = sam_plan() if mode == DEMOD_SAM else dsb_plan()
    x = busy_stream(mode, float(plan.plan.demod_rate), 12_000, 7)
    whole = run_chain(plan, x, x.size)
    for block in (1, 37, 480, 4096):
        cut = run_chain(plan, x, block)
        assert np.array_equal(cut.ring.view(np.uint32), whole.ring.view(np.uint32)), block
        assert np.array_equal(cut.audio.view(np.uint32), whole.audio.view(np.uint32)), block
    # the loop did something: locked, and the ring is not the input
    assert whole.probes[-1].locked
    assert not np.array_equal(whole.ring[:2 * x.size], x.view(np.float32))


@pytest.mark.parametrize("mode", [DEMOD_SAM, DEMOD_DSB])
def test_loops_lock_within_the_times_stated(mode):
    """SAM on the carrier within 100 ms from up to 500 Hz off; DSB within
    100 ms up to 200 Hz and within 600 ms from 500 Hz; five seeds, 30 and 10
    dB SNR, a steady tone and a twenty-tone programme."""
    sam = mode == DEMOD_SAM
    plan = sam_plan() if sam else dsb_plan()
    rate = float(plan.plan.demod_rate)
    for tone in (True, False):
        for snr in (30.0, 10.0):
            for offset in (0.0, 200.0, 500.0):
                worst, on_carrier = 0.0, 0
                for seed in range(1, 6):
                    audio = (lambda t: np.cos(TWO_PI * TONE_HZ * t)) if tone \
                        else programme(1.0, seed)
                    if sam:
                        env = (lambda a: lambda t: (1.0 + 0.8 * a(t)).astype(np.complex128))(audio)
                    else:
                        env = (lambda a: lambda t: a(t).astype(np.complex128))(audio)
                    sig = Signal(env, (lambda o: lambda t: np.full_like(t, o))(offset),
                                 0.37 * seed)
                    clean = render(sig, rate, int(rate))
                    power = float(np.mean(np.abs(clean.astype(np.complex128)) ** 2))
                    noise = math.sqrt(power / 10.0 ** (snr / 10.0))
                    x = render(sig, rate, int(rate), noise, seed * 977 + int(offset))
                    run = run_chain(plan, x, 48)
                    worst = max(worst, lock_time(run))
                    if abs(settled_frequency(run, 0.1) - offset) < 1.0:
                        on_carrier += 1
                bound = (0.1 if offset <= 500 else 0.0) if sam else \
                    (0.1 if offset <= 200 else 0.6 if offset <= 500 else 0.0)
                what = f"mode {mode} tone={tone} snr={snr} offset={offset}"
                assert worst < bound, (what, worst)
                assert on_carrier == 5, (what, on_carrier)


@pytest.mark.parametrize("mode", [DEMOD_SAM, DEMOD_DSB])
def test_loop_follows_a_drifting_carrier_to_the_passband_edge_and_no_further(mode):
    sam = mode == DEMOD_SAM
    plan = sam_plan() if sam else dsb_plan()
    rate = float(plan.plan.demod_rate)
    reach = float(max(-plan.plan.passband_low, plan.plan.passband_high))
    nco = nco_for(plan)
    drift = 100.0
    for direction in (1.0, -1.0):
        sig = Signal(am_tone(1.0, 0.8) if sam else dsb_tone(),
                     (lambda d: lambda t: d * np.maximum(0.0, t - 0.2) * drift)(direction))
        total = int((0.2 + (reach + 300.0) / drift) * rate)
        x = render(sig, rate, total, 0.01, 11)
        ring_n, block = 1 << 12, 512
        ring = np.zeros(2 * ring_n, dtype=np.float32)
        state = np.zeros(8, dtype=np.uint32)
        tracked = furthest = 0.0
        for at in range(0, total - block + 1, block):
            slot = at % ring_n
            ring[2 * slot:2 * (slot + block)] = x[at:at + block].view(np.float32)
            oracle.vrx_carrier(plan.plan.carrier, plan.carrier_block(ring_n - 1, at, block),
                               nco, ring, state)
            f = state_frequency(state, rate)
            furthest = max(furthest, abs(f))
            truth = direction * max(0.0, (at + block) / rate - 0.2) * drift
            if state[5] and abs(f - truth) < 10.0:
                tracked = max(tracked, abs(truth))
        assert tracked > 0.98 * reach, (direction, tracked, reach)
        assert furthest <= reach + 0.5, (direction, furthest, reach)


def test_dsb_comes_back_at_unit_level_whatever_the_carrier_phase():
    plan = dsb_plan()
    rate = float(plan.plan.demod_rate)
    noise = math.sqrt(0.5 / 1.0e4)
    lowest, highest, lowest_open, slowest = 1e9, 0.0, 1e9, 0.0
    for k in range(16):
        for offset in (0.0, 37.0):
            sig = Signal(dsb_tone(), (lambda o: lambda t: np.full_like(t, o))(offset),
                         TWO_PI * k / 16.0)
            x = render(sig, rate, int(0.6 * rate), noise, 100 + k)
            run = run_chain(plan, x, 48)
            slowest = max(slowest, lock_time(run))
            level = tone_amplitude(audio_between(run, rate, 0.3, 0.6), rate, TONE_HZ)
            lowest, highest = min(lowest, level), max(highest, level)
            if offset == 0.0:
                open_run = run_chain(plan, x, 480, carrier=False)
                lowest_open = min(lowest_open, tone_amplitude(
                    audio_between(open_run, rate, 0.3, 0.6), rate, TONE_HZ))
    assert lowest > 0.995 and highest < 1.005, (lowest, highest)
    assert slowest < 0.15, slowest
    assert lowest_open < 0.1, lowest_open
