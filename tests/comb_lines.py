"""Comb lines and signals for the comb kernels' tests, after
tests/reference/test_comb_kernels.cpp:28-99 (exact_tone, make_comb, noise,
wrapping_params)."""
from __future__ import annotations

import math

import numpy as np

from revenant_amd import _lib
from revenant_amd.dsp import comb_step

RING = 1 << 16
MASK = RING - 1
RATE = 2_400_000


def exact_tone(phase0: int, step: int, n: np.ndarray) -> np.ndarray:
    """The line's tone synthesised in double from the integer phase."""
    phase = (phase0 + n.astype(np.uint64) * step) % (1 << 32)
    return np.exp(2j * math.pi * phase.astype(np.float64) / 4294967296.0)


def make_comb(count: int, seed: int):
    """Lines at c + k d, random phases and amplitudes 0.02..0.2."""
    g = np.random.default_rng(seed)
    arr = (_lib.CombLine * max(count, 1))()
    c, d = 1234.5, 4567.25
    k_lo = -(count // 2)
    for l in range(count):
        arr[l].phase0 = int(g.integers(0, 1 << 32))
        arr[l].step = comb_step(c + (k_lo + l) * d, RATE)
        a = np.float32(g.uniform(0.02, 0.2)) * np.exp(1j * np.float32(g.uniform(0, 6.2831853)))
        arr[l].amplitude[0] = np.float32(a.real)
        arr[l].amplitude[1] = np.float32(a.imag)
    return arr


def noise(length: int, seed: int, sigma: float) -> np.ndarray:
    g = np.random.default_rng(seed)
    return (g.normal(0, sigma, length) + 1j * g.normal(0, sigma, length)).astype(np.complex64)


def wrapping_params(lines: int, sub_blocks: int) -> _lib.CombParams:
    """A dispatch that starts near the ring's end and crosses it."""
    return _lib.CombParams(MASK, RING - 5000, 3 * 16384, lines, sub_blocks, MASK, 0, 0)
