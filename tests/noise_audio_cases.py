"""Audio and parameters for the line and spectral kernels' tests (after
tests/reference/test_noise.cpp:126-155 and :330-397)."""
import numpy as np

from revenant_amd import _lib
from revenant_amd.dsp import (CW_GUARD_DEPTH_DB, LINE_GUARD, design_notch, line_notch_sections,
                              set_notch)

# dispatch lengths on and off every boundary: the hop, one sample, zero, a long run
LENGTHS = (100, 1, 0, 256, 700, 33, 1024, 255, 2, 3000)
RATE = 48_000


def voice_audio(count: int, seed: int, rate: float) -> np.ndarray:
    """Two steady tones, a gliding third and Gaussian noise (sigma 0.05)."""
    rng = np.random.default_rng(seed)
    t = np.arange(count) / rate
    glide = 600.0 + 200.0 * np.sin(2 * np.pi * 3.0 * t)
    v = (0.3 * np.sin(2 * np.pi * 1000.0 * t) + 0.2 * np.sin(2 * np.pi * 1700.0 * t + 0.4) +
         0.2 * np.sin(2 * np.pi * glide * t) + rng.normal(0.0, 0.05, count))
    return v.astype(np.float32)


def line_params(flags: int) -> _lib.LineParams:
    """The reference test's figures: a 1 kHz 40 dB 100 Hz notch, cascade
    sections at 2.2, 3.1 and 0.45 kHz, stride 7, delay 48, mu 0.002, and the
    CW guard at 700 Hz."""
    p = _lib.LineParams(flags=flags, stride=7, delay=48, mu=0.002,
                        leak=float(np.float32(1.0 - 1.0e-5)), eps=1.0e-6)
    set_notch(p, 0, design_notch(1000.0, 40.0, 100.0, RATE))
    more = ((2200.0, 30.0, 80.0), (3100.0, 50.0, 150.0), (450.0, 20.0, 60.0))
    for s in range(1, line_notch_sections(flags)):
        set_notch(p, s, design_notch(*more[s - 1], RATE))
    if flags & LINE_GUARD:
        set_notch(p, 0, design_notch(700.0, CW_GUARD_DEPTH_DB, 200.0, RATE), guard=True)
    return p
