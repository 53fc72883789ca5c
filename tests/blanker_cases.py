"""Channel rings for the impulse blanker's tests (after
tests/reference/test_noise.cpp:163-322): complex noise over three channels
with impulses of several heights and widths on the channel under test, and the
parameters of a span that crosses the end of both rings."""
import numpy as np

from revenant_amd import _lib
from revenant_amd.dsp import blanker_threshold, design_blanker

CHANNEL_BLOCKS = 4096
CHANNELS = 3
OUT_RING = 2048


def impulsive_channels(seed: int) -> np.ndarray:
    """CHANNELS * CHANNEL_BLOCKS complex samples uniform in [-0.05, 0.05), and
    60 impulses (height 0.2..1.4, one to three samples wide) on the channel at
    base CHANNEL_BLOCKS, placed modulo the channel length."""
    rng = np.random.default_rng(seed)
    n = CHANNEL_BLOCKS * CHANNELS
    ch = rng.uniform(-0.05, 0.05, (n, 2)).astype(np.float32)
    base = CHANNEL_BLOCKS
    for k in range(60):
        at = int(rng.integers(0, CHANNEL_BLOCKS))
        height = np.float32(0.2) * np.float32(1 + k % 7)
        for w in range(1 + k % 3):
            ch[base + ((at + w) & (CHANNEL_BLOCKS - 1))] = (height, np.float32(-0.5) * height)
    return ch


def wrapping_params(count: int = 1500, threshold_db: float = 12.0):
    """A span starting 700 samples before the end of the third block: it
    crosses the end of both the channel and the out ring."""
    first = 3 * CHANNEL_BLOCKS - 700
    p = _lib.BlankerParams()
    p.chan_base = CHANNEL_BLOCKS
    p.chan_mask = CHANNEL_BLOCKS - 1
    p.chan_first = first & p.chan_mask
    p.out_mask = OUT_RING - 1
    p.out_first = first & p.out_mask
    p.count = count
    p.threshold = blanker_threshold(threshold_db)
    return p, first


def apply_params(p, first: int, config):
    """The apply span over the part the detect span's flags cover."""
    a = _lib.BlankerParams.from_buffer_copy(bytes(p))
    a.chan_first = (first + config.hang) & p.chan_mask
    a.out_first = (first + config.hang) & p.out_mask
    a.count = p.count - config.hang - config.lead
    return a


def config_144k():
    return design_blanker(144_000)
