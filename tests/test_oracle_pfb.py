"""Pins the channelizer: design tables, the radix-2 transform and the branch
filter + transform composition, each against something other than itself.

After tests/reference/test_pfb.cpp:155-513.
"""
import math

import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import GridParams, build_twiddles, design_prototype


@pytest.mark.parametrize("M", [2, 4, 8, 64, 256, 4096, 16384])
def test_twiddles_closed_form(M):
    w = build_twiddles(M).astype(np.complex128)
    exact = np.exp(-2j * np.pi * np.arange(M) / M)
    assert np.max(np.abs(w - exact)) < 1.2e-7
    assert np.max(np.abs(np.abs(w) - 1.0)) < 1.2e-7
    f = build_twiddles(M).view(np.float32)
    assert not np.any((f == 0) & np.signbit(f)), "no -0.0 in the table"
    if M >= 4:
        want = {0: (1, 0), M // 4: (0, -1), M // 2: (-1, 0), 3 * M // 4: (0, 1)}
    else:
        want = {0: (1, 0), 1: (-1, 0)}
    for j in want:
        assert (w[j].real, w[j].imag) == want[j], j


@pytest.mark.parametrize("M", [4, 8, 64, 512, 2048, 4096, 16384])
def test_twiddles_are_exactly_quadrant_symmetric(M):
    """Entry k + M/4 is entry k turned by -i, (x, y) -> (y, -x), bit for bit,
    for k below M/4: the chain kernel reads only the lower half of each
    stage's twiddles and forms the upper half this way."""
    t = build_twiddles(M).view(np.float32).reshape(M, 2)
    lo, hi = t[:M // 4], t[M // 4:M // 2]
    assert np.array_equal(hi[:, 0].view(np.uint32), lo[:, 1].view(np.uint32))
    assert np.array_equal(hi[:, 1].view(np.uint32), (-lo[:, 0]).view(np.uint32))


@pytest.mark.parametrize("M,L1", [(64, 17), (64, 16), (4096, 16), (8, 5)])
def test_prototype_invariants(M, L1):
    g = GridParams(M, L1, M // 2)
    p = design_prototype(g)
    h = p.taps.astype(np.float64)
    used = M * (L1 - 1) + 1
    assert h.size == M * L1
    assert np.all(h[used:] == 0.0)                            # zero pad
    assert abs(h.sum() - 1.0) < 1e-6                          # unit DC gain
    assert np.array_equal(h[:used], h[:used][::-1])           # symmetric bit for bit
    assert p.group_delay_samples == M * (L1 - 1) // 2


def test_prototype_stopband_at_120db():
    """test_pfb.cpp:155-187: M=64, 17 taps/branch, 120 dB -> below -110 dB."""
    p = design_prototype(GridParams(64, 17, 32), 120.0)
    assert p.stopband_db < -110.0
    # independent check of the stopband on a dense direct evaluation
    h = p.taps.astype(np.float64)
    f = np.linspace(0.75 / 64, 0.5, 4000)
    resp = np.abs(np.exp(-2j * np.pi * np.outer(f, np.arange(h.size))) @ h)
    assert 20 * np.log10(resp.max()) < -110.0


def test_prototype_beta_override():
    """Kaiser beta 8 (the workload's prototype) is the window I0(8 sqrt(1-x^2))/I0(8)."""
    M, L1 = 16, 5
    p = design_prototype(GridParams(M, L1, 8), beta=8.0)
    n = np.arange(M * (L1 - 1) + 1)
    c = 0.5 * (n.size - 1)
    x = (n - c) / c
    w = np.i0(8.0 * np.sqrt(1 - x ** 2)) / np.i0(8.0)
    s = np.sinc((n - c) / M) / M
    h = s * w
    h /= h.sum()
    assert np.max(np.abs(p.taps[:n.size] - h)) < 1e-7


@pytest.mark.parametrize("size", [2, 4, 8, 16, 32, 64, 128, 256, 1024, 4096])
def test_fft_against_fp64_dft(size):
    tw = build_twiddles(size)
    x = np.random.default_rng(size).uniform(-1, 1, 2 * size).astype(np.float32)
    y = oracle.fft_radix2(tw, x, size, 1).view(np.complex64).astype(np.complex128)
    ref = oracle.fp64_dft(x.view(np.complex64))
    assert np.max(np.abs(y - ref)) <= 4e-7 * math.log2(size) * math.sqrt(size)


def test_fft_impulse_is_exactly_flat():
    for size in (8, 64, 4096):
        x = np.zeros(2 * size, np.float32)
        x[0] = 1.0
        y = oracle.fft_radix2(build_twiddles(size), x, size, 1).view(np.complex64)
        assert np.all(y == np.complex64(1.0))


def _channelize(grid, x, first, blocks, taps):
    M, D = grid.channels, grid.decimation
    cap = 1 << (x.size + 1).bit_length()
    ring = np.zeros(2 * cap, np.float32)
    ring[:2 * x.size] = x.astype(np.complex64).view(np.float32)
    br = oracle.pfb_branch(grid.c(), taps, ring, _lib.PfbBranchParams(cap - 1, first * D, blocks))
    rb = 1 << (blocks - 1).bit_length()
    out = np.zeros(2 * M * rb, np.float32)
    oracle.pfb_fft(_lib.PfbFftParams(M, D, grid.stages, first, blocks, rb, rb - 1), br,
                   build_twiddles(M), out)
    out = out.view(np.complex64).reshape(M, rb)
    return np.stack([out[:, (first + b) & (rb - 1)] for b in range(blocks)])


@pytest.mark.parametrize("M,L1,D", [(8, 4, 8), (8, 4, 4), (64, 17, 32), (64, 16, 64),
                                    (256, 16, 128)])
def test_channelizer_against_fp64_definition(M, L1, D):
    """y_k[m] = sum_n h[n] x[mD-n] exp(-j2pi k(mD-n)/M) (pfb_branch.comp:9-19).
    Reversal-correct is ~1e-7; a mirrored spectrum (reversal wrong) is O(1)."""
    g = GridParams(M, L1, D)
    taps = design_prototype(g).taps
    first, blocks = (M * L1 + D - 1) // D + 1, 12
    S = (first + blocks) * D + 1
    r = np.random.default_rng(M + D)
    x = (r.uniform(-1, 1, S) + 1j * r.uniform(-1, 1, S))
    x32 = x.astype(np.complex64)
    got = _channelize(g, x32, first, blocks, taps)
    ref = oracle.fp64_channelize(x32, taps.astype(np.float64), M, D, first, blocks)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 2e-6
    mirrored = ref[:, (-np.arange(M)) % M]
    assert np.max(np.abs(got - mirrored)) / np.max(np.abs(ref)) > 0.05


@pytest.mark.parametrize("k", [1, 3, 7, 16, 61, 63])
def test_tone_lands_in_its_channel(k):
    """test_pfb.cpp:429-513: a tone at channel k's centre reads 1 +- 0.02 there,
    below 1e-3 in every non-adjacent channel."""
    M, L1, D = 64, 17, 32
    g = GridParams(M, L1, D)
    taps = design_prototype(g).taps
    first, blocks = (M * L1) // D + 1, 8
    n = np.arange((first + blocks) * D + 1)
    x = np.exp(2j * np.pi * k * n / M)
    y = np.abs(_channelize(g, x.astype(np.complex64), first, blocks, taps))
    assert np.all(np.abs(y[:, k] - 1.0) < 0.02)
    others = np.delete(y, [k, (k - 1) % M, (k + 1) % M], axis=1)
    assert others.max() < 1e-3
