"""Spectra built bin by bin for the comb detector's cases
(tests/detect/test_comb_detect.cpp:28-147): a floor of gamma-distributed bins
(K frames averaged, each bin exponential), lines drawn with a Hann window's
power response at fractional bin positions, wideband plateaus. Seeded."""
import numpy as np

BINS = 16384
RATE = 2_400_000
HZ_PER_BIN = RATE / BINS


def hann_power(u: np.ndarray) -> np.ndarray:
    """A Hann window's power response u bins from a line, 1 at u = 0."""
    u = np.asarray(u, dtype=np.float64)

    def dirichlet(v):
        out = np.ones_like(v)
        nz = np.abs(v) > 1e-9
        out[nz] = np.sin(np.pi * v[nz]) / (np.pi * v[nz])
        return out

    w = 0.5 * dirichlet(u) + 0.25 * (dirichlet(u - 1.0) + dirichlet(u + 1.0))
    return w * w / 0.25


class Scene:
    def __init__(self, seed: int, frames: float = 140.0):
        self.g = np.random.default_rng(seed)
        self.frames = frames
        self.mean = np.ones(BINS)

    def line(self, x: float, height_db: float) -> None:
        a = 10.0 ** (height_db / 10.0)
        c = int(round(x))
        i = np.arange(max(0, c - 40), min(BINS, c + 41))
        self.mean[i] += a * hann_power(i - x)

    def plateau(self, start: int, width: int, height_db: float) -> None:
        i = np.arange(start, min(BINS, start + width))
        db = height_db + self.g.uniform(-1.0, 1.0, i.size)
        self.mean[i] += 10.0 ** (db / 10.0)

    def comb(self, c: float, d: float, lo_db: float, hi_db: float, keep: float = 1.0) -> int:
        k_lo = int(np.ceil((2.0 - c) / d))
        k_hi = int(np.floor((BINS - 3.0 - c) / d))
        drawn = 0
        for k in range(k_lo, k_hi + 1):
            h = self.g.uniform(lo_db, hi_db)
            if self.g.random() < keep:
                self.line(c + k * d, h)
                drawn += 1
        return drawn

    def render(self) -> np.ndarray:
        return (self.mean * self.g.gamma(self.frames, 1.0 / self.frames, BINS)).astype(np.float32)


def line_nearest_dc(c: float, d: float) -> float:
    """The comb's line nearest DC in hertz, for a comb drawn at c + k d bins."""
    dc = BINS // 2
    k = round((dc - c) / d)
    return (c + k * d - dc) * HZ_PER_BIN


def on_comb(found, c: float, d: float, hz: float) -> bool:
    """`found` lies on the comb c + k d bins with its spacing within `hz`."""
    spacing = d * HZ_PER_BIN
    offset = (c - BINS // 2) * HZ_PER_BIN
    k = round((found.offset_hz - offset) / spacing)
    return abs(found.spacing_hz - spacing) < hz and \
        abs(found.offset_hz - (offset + k * spacing)) < 0.2 * HZ_PER_BIN
