"""Pins the spectrum, the levels and the averaged power spectrum.

After tests/reference/test_spectrum.cpp:419-1045 and
test_spectrum_levels.cpp:164-321; the averaged spectrum has no reference twin
and is pinned against its FP64 definition.
"""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import build_spectrum_window, build_twiddles, ema_params


def test_window_unit_coherent_gain():
    for n in (16, 256, 2048, 4096):
        w = build_spectrum_window(n)
        assert abs(w[:n].astype(np.float64).sum() - 1.0) < 1e-6
        gains = w[n:]
        assert np.all(gains >= 1.0 - 1e-6) and np.all(gains <= 10 ** 1.2 * (1 + 1e-6))


def test_tone_lands_in_its_bin():
    """A complex tone in one channel peaks at the matching kept bin, at 0 dB
    after the correction when it sits mid-channel."""
    M, N = 8, 256
    blocks = N
    ring = np.zeros((M, blocks), np.complex64)
    k_bin = 37  # output bin index within the kept half
    f = (k_bin - N / 4) / N  # cycles per channel sample
    ring[3] = np.exp(2j * np.pi * f * np.arange(blocks)).astype(np.complex64)
    p = _lib.SpectrumParams(M, N, 8, blocks, blocks - 1, 0)
    frame = oracle.spectrum(p, ring.view(np.float32).ravel(), build_twiddles(N),
                            build_spectrum_window(N))
    row = frame.reshape(M, N // 2)[(3 + M // 2) % M]
    assert int(np.argmax(row)) == k_bin
    w = build_spectrum_window(N)
    expected_db = 10 * np.log10(float(w[N + k_bin]))
    assert abs(row[k_bin] - expected_db) < 0.01
    # every other channel sits on the -200 dB floor
    assert np.all(np.delete(frame.reshape(M, N // 2), (3 + M // 2) % M, axis=0) == np.float32(
        oracle.det("log2", np.float32([1e-20]))[0] * np.float32(3.01029995663981195)))


@pytest.mark.parametrize("bins", [1, 10, 999, 65536])
def test_levels_against_exact_sort(bins):
    g = np.random.default_rng(bins)
    frame = g.uniform(-180, 20, bins).astype(np.float32)
    lv = oracle.spectrum_levels(_lib.SpectrumLevelsParams(bins, 50, 990), frame)
    s = np.sort(frame)
    for v, pm in zip(lv, (50, 990)):
        rank = min((bins // 1000) * pm + ((bins % 1000) * pm) // 1000, bins - 1)
        assert abs(v - s[rank]) <= 0.234375 / 2 + 1e-4


def test_levels_bucket_boundaries():
    frame = np.float32([-200.0, -199.8, 39.9, 100.0, -1000.0])
    lv = oracle.spectrum_levels(_lib.SpectrumLevelsParams(5, 0, 1000), frame)
    assert lv[0] == np.float32(-200.0 + 0.5 * 0.234375)
    assert lv[1] == np.float32(-200.0 + 1023.5 * 0.234375)


@pytest.mark.parametrize("F,G", [(16, 1), (4, 4), (2, 8), (1, 16)])
def test_ema_against_fp64_definition(F, G):
    """Frame partial + row partial + recurrence equals the FP64 exponential
    average of the same powers to float rounding, for any frame and row size;
    a mistaken decay factor or a dropped term fails by orders of magnitude."""
    M, frames = 16, 640 // F
    blocks = 1024
    g = np.random.default_rng(3)
    ring = (g.standard_normal((M, blocks)) + 1j * g.standard_normal((M, blocks))).astype(
        np.complex64)
    p = ema_params(M, blocks, 0, F, frames, 1 / 16, G)
    part = oracle.ema_partial(p, ring.view(np.float32).ravel())
    state = np.zeros(M, np.float32)
    rows = oracle.ema_combine(p, part, state).reshape(frames // G, M)
    x = np.abs(ring.astype(np.complex128)) ** 2
    P = np.zeros(M)
    for r in range(frames // G):
        for i in range(F * G):
            P = P + (x[:, r * F * G + i] - P) / 16.0
        assert np.max(np.abs(rows[r] - P) / P) < 3e-6
    assert np.array_equal(state, rows[-1])


def test_ema_alpha_one_and_constant_input():
    M, F = 4, 8
    ring = np.full((M, 64), 2.0 + 0j, np.complex64)
    ring[:, 7] = 3.0
    p = ema_params(M, 64, 0, F, 1, 1.0)
    assert np.all(oracle.ema_partial(p, ring.view(np.float32).ravel()) == 9.0)
    ring[:, 7] = 2.0
    alpha = np.float32(1 / 16)
    p = ema_params(M, 64, 0, F, 8, alpha, 2)
    part = oracle.ema_partial(p, ring.view(np.float32).ravel())
    rows = oracle.ema_combine(p, part, np.zeros(M, np.float32)).reshape(4, M)
    assert abs(rows[-1, 0] - 4.0 * (1 - (15 / 16) ** 64)) < 1e-5
