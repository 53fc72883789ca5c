"""Signals at the demodulation rate and the carrier -> detector chain on the
oracle, after tests/reference/test_vrx_carrier.cpp:90-300 (Signal, Synth,
programme, am_tone, dsb_tone, run_chain, lock_time, settled_frequency).

Synthesis is numpy in float64, rounded to complex64 once. The phase is
accumulated from the offset, so a frequency that moves is a frequency and
not a phase jump.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (CARRIER_NONE, DEMOD_DSB, DEMOD_SAM, GridParams, VrxRequest,
                              build_nco_table, plan_vrx)

TWO_PI = 2.0 * math.pi
SOURCE_RATE = 2_400_000
GRID = GridParams(64, 17, 32)  # tests/reference/test_vrx_carrier.cpp:62-67
TONE_HZ = 1000.0


def make_plan(mode: int, bandwidth: int):
    return plan_vrx(GRID, SOURCE_RATE, VrxRequest(196_500, mode, bandwidth))


def sam_plan():
    return make_plan(DEMOD_SAM, 10_000)


def dsb_plan():
    return make_plan(DEMOD_DSB, 6_000)


_NCO = {}


def nco_for(plan) -> np.ndarray:
    k = plan.plan.carrier.nco_log2
    if k not in _NCO:
        _NCO[k] = build_nco_table(k)
    return _NCO[k]


@dataclass
class Signal:
    envelope: Callable[[np.ndarray], np.ndarray]           # t -> complex envelope
    offset_hz: Callable[[np.ndarray], np.ndarray] = lambda t: np.zeros_like(t)
    phase: float = 0.0


def render(sig: Signal, rate: float, count: int, noise_rms: float = 0.0, seed: int = 1,
           start: int = 0) -> np.ndarray:
    """Samples start .. start+count-1 of the signal (complex64)."""
    n = np.arange(start, start + count, dtype=np.float64)
    t = n / rate
    if start == 0:
        steps = TWO_PI * sig.offset_hz(t) / rate
        ph = sig.phase + np.concatenate([[0.0], np.cumsum(steps[:-1])])
    else:  # a later piece: accumulate from the beginning
        full = np.arange(0, start + count, dtype=np.float64) / rate
        steps = TWO_PI * sig.offset_hz(full) / rate
        ph = (sig.phase + np.concatenate([[0.0], np.cumsum(steps[:-1])]))[start:]
    z = sig.envelope(t) * np.exp(1j * np.mod(ph, TWO_PI))
    if noise_rms > 0.0:
        g = np.random.default_rng(seed)
        z = z + (g.standard_normal(count) + 1j * g.standard_normal(count)) * (noise_rms / math.sqrt(2))
    return z.astype(np.complex64)


def programme(depth: float, seed: int) -> Callable[[np.ndarray], np.ndarray]:
    """Twenty tones at fixed random frequencies (300..3000 Hz) and phases,
    scaled to a peak of `depth`: a stand-in for speech with no line standing
    above the others."""
    g = np.random.default_rng(seed)
    hz = g.uniform(300.0, 3000.0, 20)
    ph = g.uniform(0.0, TWO_PI, 20)
    scale = depth / 20.0

    def f(t):
        return scale * np.cos(TWO_PI * hz[None, :] * t[:, None] + ph[None, :]).sum(axis=1)
    return f


def am_tone(carrier: float, m: float):
    return lambda t: (carrier + m * np.cos(TWO_PI * TONE_HZ * t)).astype(np.complex128)


def dsb_tone():
    return lambda t: np.cos(TWO_PI * TONE_HZ * t).astype(np.complex128)


def busy_stream(mode: int, rate: float, count: int, seed: int) -> np.ndarray:
    """Noise, a carrier off the centre, a fade and a step in frequency
    (test_vrx_carrier.cpp:420-437)."""
    if mode == DEMOD_SAM:
        def env(t):
            carrier = np.where((t > 0.05) & (t < 0.07), 0.05, 0.5)
            return (carrier + 0.4 * np.cos(TWO_PI * TONE_HZ * t)).astype(np.complex128)
    else:
        def env(t):
            return (0.5 * np.cos(TWO_PI * 730.0 * t)).astype(np.complex128)
    sig = Signal(env, lambda t: np.where(t < 0.08, 310.0, 190.0), 0.9)
    return render(sig, rate, count, 0.05, seed)


@dataclass
class Probe:
    at_seconds: float
    locked: bool
    frequency_hz: float


@dataclass
class Run:
    audio: np.ndarray
    history: int
    probes: list = field(default_factory=list)
    ring: np.ndarray | None = None


def state_frequency(state: np.ndarray, rate: float) -> float:
    return float(state[1:2].view(np.float32)[0]) * rate / TWO_PI


def run_chain(plan, x: np.ndarray, block: int, events=(), carrier: bool = True) -> Run:
    """The carrier loop in dispatches of `block` (events: (at, reset,
    moved_hz) retunes), then the detector over the whole stream."""
    cap = 1 << max(1, (len(x) - 1).bit_length())
    mask = cap - 1
    ring = np.zeros(2 * cap, dtype=np.float32)
    ring[:2 * len(x)] = x.view(np.float32)
    rate = float(plan.plan.demod_rate)
    probes = []
    if carrier and plan.plan.carrier.power != CARRIER_NONE:
        nco = nco_for(plan)
        state = np.zeros(8, dtype=np.uint32)
        events = list(events)
        start = 0
        while start < len(x):
            end = min(len(x), start + block)
            reset, shift = False, 0.0
            if events:
                at, ev_reset, moved = events[0]
                if at == start:
                    reset = ev_reset
                    if not reset:
                        shift = -TWO_PI * moved / rate
                    events.pop(0)
                elif at < end:
                    end = at
            p = plan.carrier_block(mask, start, end - start, reset, float(np.float32(shift)))
            oracle.vrx_carrier(plan.plan.carrier, p, nco, ring, state)
            probes.append(Probe(end / rate, bool(state[5]), state_frequency(state, rate)))
            start = end
    hist = plan.history
    dp = _lib.VrxDemodParams(mask, hist, len(x) - hist, plan.plan.demod_gain)
    audio = oracle.vrx_demod(plan.demod, dp, ring, plan.weights)
    return Run(audio, hist, probes, ring)


def lock_time(run: Run) -> float:
    """When the lock test last went true and stayed true (inf if not locked
    at the end)."""
    if not run.probes or not run.probes[-1].locked:
        return math.inf
    since = 0.0
    for p in run.probes:
        if not p.locked:
            since = p.at_seconds
    return since


def settled_frequency(run: Run, seconds: float) -> float:
    end = run.probes[-1].at_seconds
    f = [p.frequency_hz for p in run.probes if p.at_seconds > end - seconds]
    return sum(f) / max(1, len(f))


def audio_between(run: Run, rate: float, t0: float, t1: float) -> np.ndarray:
    a = int(t0 * rate) - run.history
    b = min(run.audio.size, int(t1 * rate) - run.history)
    return run.audio[a:b]


def tone_amplitude(audio: np.ndarray, rate: float, hz: float) -> float:
    """Amplitude of a tone by projection onto cos/sin over the window."""
    n = np.arange(audio.size, dtype=np.float64)
    w = TWO_PI * hz * n / rate
    a = audio.astype(np.float64)
    c = 2.0 * np.mean(a * np.cos(w))
    s = 2.0 * np.mean(a * np.sin(w))
    return math.hypot(c, s)
