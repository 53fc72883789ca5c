"""GPU vs oracle, bit for bit: full-span spectrum, levels, averaged power.

After tests/reference/test_spectrum.cpp and test_spectrum_levels.cpp.
"""
import numpy as np
import pytest

from helpers import SEED_SPECTRUM, adversarial, assert_bit_exact, from_dev, rng, to_dev, \
    uniform_complex
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import build_spectrum_window, build_twiddles, ema_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("M,N,blocks", [(64, 2048, 4096), (8, 16, 64), (16, 4, 8),
                                        (4096, 256, 512), (32, 4096, 4096)])
def test_spectrum_matches_twin(dev, M, N, blocks):
    ring = uniform_complex(SEED_SPECTRUM + N, M * blocks, 0.5)
    tw = build_twiddles(N)
    win = build_spectrum_window(N)
    p = _lib.SpectrumParams(M, N, N.bit_length() - 1, blocks, blocks - 1, blocks - 5)
    cpu = oracle.spectrum(p, ring, tw, win)
    frame = torch.empty(M * N // 2, dtype=torch.float32, device="cuda")
    dev.spectrum(p, to_dev(ring), to_dev(tw.view(np.float32)), to_dev(win), frame)
    assert_bit_exact(from_dev(frame), cpu, SEED_SPECTRUM, f"spectrum M={M} N={N}")


@pytest.mark.parametrize("bins", [1, 7, 1000, 65536, 100003])
def test_levels_match_twin(dev, bins):
    g = rng(SEED_SPECTRUM + bins)
    frame = g.uniform(-230.0, 60.0, size=bins).astype(np.float32)
    frame[: bins // 10] = -200.0  # a floor plateau, many ties
    p = _lib.SpectrumLevelsParams(bins, 50, 990)
    cpu = oracle.spectrum_levels(p, frame)
    out = torch.zeros(2, dtype=torch.float32, device="cuda")
    dev.spectrum_levels(p, to_dev(frame), out)
    assert_bit_exact(from_dev(out), cpu, what=f"levels {bins}")


@pytest.mark.parametrize("M,frame_blocks,frames,row_frames",
                         [(64, 16, 32, 1), (4096, 4, 64, 4), (5, 3, 7, 7), (33, 2, 60, 3),
                          (64, 2, 200, 2), (8, 1, 1000, 10)])
def test_ema_matches_twin(dev, M, frame_blocks, frames, row_frames):
    blocks = 1 << (frame_blocks * frames - 1).bit_length()
    ring = adversarial(SEED_SPECTRUM + M, M * blocks)
    ring[np.abs(ring) > 1e18] = 0.5  # keep |x|^2 finite so the recurrence is exercised
    p = ema_params(M, blocks, 3, frame_blocks, frames, 1.0 / 16.0, row_frames)
    part_cpu = oracle.ema_partial(p, ring)
    state_cpu = np.linspace(0.0, 1.0, M, dtype=np.float32)
    rows_cpu = oracle.ema_combine(p, part_cpu, state_cpu)
    part = torch.zeros(frames * M, dtype=torch.float32, device="cuda")
    dev.ema_partial(p, to_dev(ring), part)
    assert_bit_exact(from_dev(part), part_cpu, what="ema partial")
    state = to_dev(np.linspace(0.0, 1.0, M, dtype=np.float32))
    rows = torch.zeros((frames // row_frames) * M, dtype=torch.float32, device="cuda")
    dev.ema_combine(p, part, state, rows)
    assert_bit_exact(from_dev(rows), rows_cpu, what="ema rows")
    assert_bit_exact(from_dev(state), state_cpu, what="ema state")
