"""The adaptive noise profile (rvx_noise_profile_*) against what
core/dsp/noise_profile.h promises, with the cases of the reference's
tests/reference/test_noise_profile.cpp (unit reconstruction at zero strength,
hum cut by about the strength once learned, a tone above the noise kept within
1 dB, the quiet average held through an open passage, minimum statistics with
no squelch, a reset dropping a learned profile) and pins of this build's own:
the noise floor's depth in closed form, the transform size per rate, the
pass-through when off, retune and rate change. Host only: no GPU."""
import math

import numpy as np
import pytest

from revenant_amd.noise_profile import NoiseProfile, NoiseProfileSettings

RATE = 48_000.0
HUM1, HUM3, TONE = 120.0, 360.0, 1_500.0


class Source:
    """Hum at 120 and 360 Hz over white noise (sigma 0.01), an optional tone."""

    def __init__(self, seed=20261020):
        self.rng = np.random.default_rng(seed)
        self.n = 0

    def next(self, count, tone):
        t = (self.n + np.arange(count)) / RATE
        self.n += count
        v = (0.05 * np.sin(2 * np.pi * HUM1 * t) + 0.03 * np.sin(2 * np.pi * HUM3 * t + 0.4)
             + self.rng.normal(0.0, 0.01, count))
        if tone:
            v = v + 0.1 * np.sin(2 * np.pi * TONE * t)
        return v.astype(np.float32)


def tone_power(x, hz):
    i = np.arange(len(x))
    acc = np.sum(x.astype(np.float64) * np.exp(-2j * np.pi * hz * i / RATE))
    a = 2.0 * abs(acc) / len(x)
    return a * a


def db(r):
    return 10.0 * math.log10(r)


def feed(profile, src, seconds, keep, tone, quiet, settings, rate=RATE, block=733):
    """Feeds `seconds` in blocks of an awkward length; returns the input and
    output of the last `keep` seconds."""
    total = int(seconds * rate)
    start = total - int(keep * rate)
    ins, outs, done = [], [], 0
    while done < total:
        n = min(block, total - done)
        chunk = src.next(n, tone)
        original = chunk.copy()
        profile.process(chunk, rate, quiet, settings)
        lo = max(0, start - done)
        if lo < n:
            ins.append(original[lo:])
            outs.append(chunk[lo:])
        done += n
    return np.concatenate(ins), np.concatenate(outs)


def test_zero_strength_reconstructs_its_input():
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, strength_db=0.0)
    x, y = feed(p, src, 1.0, 1.0, True, False, s)
    lag = p.latency()
    assert lag == 1024
    worst = np.max(np.abs(y[2 * lag:].astype(np.float64) - x[lag:len(x) - lag]))
    assert worst < 1e-6


@pytest.mark.parametrize("strength", [6.0, 12.0, 20.0])
def test_learn_now_cuts_hum_by_the_strength_and_keeps_the_tone(strength):
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, strength_db=strength, auto_learn=False)
    feed(p, src, 0.1, 0.1, False, False, s)
    s.learn_count = 1
    feed(p, src, 1.5, 0.1, False, False, s)
    assert p.learned() and not p.learning()
    x, y = feed(p, src, 3.0, 2.0, True, False, s)
    hum1 = db(tone_power(y, HUM1) / tone_power(x, HUM1))
    hum3 = db(tone_power(y, HUM3) / tone_power(x, HUM3))
    tone = db(tone_power(y, TONE) / tone_power(x, TONE))
    assert -(strength + 3.0) < hum1 < -(strength - 3.0), hum1
    assert -(strength + 3.0) < hum3 < -(strength - 3.0), hum3
    assert abs(tone) < 1.0, tone


def test_auto_learn_from_a_shut_squelch_holds_through_the_open_passage():
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, strength_db=12.0)
    feed(p, src, 1.5, 0.1, False, True, s)
    x, y = feed(p, src, 4.0, 2.0, True, False, s)
    hum1 = db(tone_power(y, HUM1) / tone_power(x, HUM1))
    tone = db(tone_power(y, TONE) / tone_power(x, TONE))
    assert hum1 < -9.0, hum1
    assert abs(tone) < 1.0, tone


def test_minimum_statistics_finds_hum_with_no_squelch():
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, strength_db=12.0)
    x, y = feed(p, src, 5.0, 2.0, False, False, s)
    hum1 = db(tone_power(y, HUM1) / tone_power(x, HUM1))
    assert -15.0 < hum1 < -9.0, hum1


def test_reset_drops_a_learned_profile():
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, auto_learn=False)
    feed(p, src, 0.1, 0.1, False, False, s)
    s.learn_count = 1
    feed(p, src, 1.5, 0.1, False, False, s)
    assert p.learned() and p.profile().size == 513
    s.reset_count = 1
    feed(p, src, 0.1, 0.1, False, False, s)
    assert not p.learned()
    assert p.profile().size == 0


@pytest.mark.parametrize("strength", [6.0, 12.0])
def test_noise_alone_is_cut_by_exactly_the_floor(strength):
    """White noise against a profile learned from the same noise: every bin's
    smoothed power sits near its estimate, so the Wiener gain is at the floor
    10^(-strength/20) nearly everywhere and the output power is the input's
    times floor^2, -strength dB (the odd bin whose power strays above the
    estimate lifts it a little)."""
    p = NoiseProfile()
    rng = np.random.default_rng(7)
    s = NoiseProfileSettings(enabled=True, strength_db=strength, auto_learn=False)
    noise = lambda n: rng.normal(0.0, 0.05, n).astype(np.float32)
    x = noise(int(0.1 * RATE)); p.process(x, RATE, False, s)
    s.learn_count = 1
    x = noise(int(1.5 * RATE)); p.process(x, RATE, False, s)
    assert p.learned()
    x = noise(int(2.0 * RATE)); y = x.copy(); p.process(y, RATE, False, s)
    lag = p.latency()
    ratio = db(np.mean(y[lag:].astype(np.float64) ** 2) / np.mean(x.astype(np.float64) ** 2))
    assert -strength - 0.5 < ratio < -strength + 2.0, ratio


def test_transform_size_by_rate_and_off_is_a_pass_through():
    for rate, size in ((48_000.0, 1024), (32_000.0, 1024), (24_000.0, 512), (16_000.0, 512),
                       (8_000.0, 256)):
        p = NoiseProfile()
        x = Source().next(3000, True)
        y = x.copy()
        p.process(y, rate, False, NoiseProfileSettings(enabled=False))
        assert np.array_equal(x, y)  # off: untouched
        p.process(y, rate, False, NoiseProfileSettings(enabled=True, strength_db=0.0))
        assert p.latency() == size
        # the first `size` samples out are the zero history the stream starts
        # with (to the transform's round-off), then the input
        assert np.max(np.abs(y[:size])) < 1e-12 and np.max(np.abs(y[size:])) > 1e-3


def test_retune_keeps_a_learned_profile_and_a_rate_change_drops_it():
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, auto_learn=False, learn_count=0)
    feed(p, src, 0.1, 0.1, False, False, s)
    s.learn_count = 1
    feed(p, src, 1.5, 0.1, False, False, s)
    learned = p.profile().copy()
    p.retune()
    assert p.learned()
    feed(p, src, 0.1, 0.1, False, False, s)
    assert np.array_equal(p.profile(), learned)
    feed(p, src, 0.1, 0.1, False, False, s, rate=24_000.0)
    assert not p.learned() and p.latency() == 512


def test_quiet_average_wins_over_minimum_statistics():
    """After a shut-squelch passage the estimate is the quiet average, not the
    minimum (with a steady carrier in the open passage, minimum statistics
    would take the carrier for noise)."""
    p, src = NoiseProfile(), Source()
    s = NoiseProfileSettings(enabled=True, strength_db=12.0)
    feed(p, src, 1.0, 0.1, False, True, s)
    quiet = p.profile().copy()
    assert quiet.size == 513
    feed(p, src, 3.0, 0.1, True, False, s)
    # the quiet average is held through the open passage: unchanged
    assert np.array_equal(p.profile(), quiet)


def test_strength_is_clamped_and_checked():
    p = NoiseProfile()
    x = Source().next(4096, False)
    p.process(x.copy(), RATE, False, NoiseProfileSettings(enabled=True, strength_db=60.0))
    with pytest.raises(RuntimeError):
        p.process(x.copy(), RATE, False, NoiseProfileSettings(enabled=True,
                                                              strength_db=float("nan")))
