"""Seeded synthetic IQ scenes (the benchmark's and the tests' inputs).

Modelled on core/dsp/synth (wideband.h, channel.h): carriers are pure
functions of the absolute sample index, with the carrier phase reduced
exactly in integers (f * n mod rate) before it becomes an angle, so a sample
at index 2^27 is as accurate as one at index 0. Noise is seeded complex
Gaussian. Works on any torch device; the benchmark builds its 2^27-sample
scene on the GPU in chunks.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

KINDS = ("cw", "am", "nfm", "wfm")


@dataclass
class Carrier:
    kind: str
    freq_hz: int          # offset from the centre, integer hertz
    level_dbfs: float
    tone_hz: int = 1000   # modulating tone
    deviation_hz: int = 0 # FM deviation
    depth: float = 0.5    # AM depth


def wideband_scene(rate: int, count: int, seed: int, level_range=(-80.0, -10.0),
                   guard: float = 0.45) -> list[Carrier]:
    """`count` carriers of mixed kinds at uniform frequencies in +-guard*rate."""
    g = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = KINDS[i % 4]
        f = int(g.integers(-int(guard * rate), int(guard * rate)))
        lvl = float(g.uniform(*level_range))
        if kind == "nfm":
            out.append(Carrier(kind, f, lvl, int(g.integers(300, 3000)), 2500))
        elif kind == "wfm":
            out.append(Carrier(kind, f, lvl, int(g.integers(300, 15000)), 75000))
        elif kind == "am":
            out.append(Carrier(kind, f, lvl, int(g.integers(300, 5000)), 0, 0.8))
        else:
            out.append(Carrier(kind, f, lvl))
    return out


def _turns(freq_hz: int, n: torch.Tensor, rate: int) -> torch.Tensor:
    """Exact phase in turns of a carrier at freq_hz: ((f * n) mod rate) / rate."""
    return torch.remainder(n * int(freq_hz), rate).to(torch.float64) / rate


def render(carriers: list[Carrier], rate: int, start: int, count: int, noise_dbfs: float | None,
           seed: int, device="cpu", chunk: int = 1 << 22) -> torch.Tensor:
    """complex64 samples start .. start+count of the scene (full scale = 1.0)."""
    out = torch.empty(count, dtype=torch.complex64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    two_pi = 2.0 * math.pi
    for c0 in range(0, count, chunk):
        c1 = min(count, c0 + chunk)
        n = torch.arange(start + c0, start + c1, dtype=torch.int64, device=device)
        acc = torch.zeros(c1 - c0, dtype=torch.complex128, device=device)
        for c in carriers:
            amp = 10.0 ** (c.level_dbfs / 20.0)
            ph = _turns(c.freq_hz, n, rate) * two_pi
            mod = _turns(c.tone_hz, n, rate) * two_pi
            if c.kind in ("nfm", "wfm"):
                ph = ph + (c.deviation_hz / c.tone_hz) * torch.sin(mod)
                env = amp
            elif c.kind == "am":
                env = amp * (1.0 + c.depth * torch.cos(mod)) / (1.0 + c.depth)
            else:
                env = amp
            acc += env * torch.polar(torch.ones_like(ph), ph)
        if noise_dbfs is not None:
            sigma = 10.0 ** (noise_dbfs / 20.0) / math.sqrt(2.0)
            noise = torch.randn(2, c1 - c0, generator=gen, device=device, dtype=torch.float64)
            acc += torch.complex(noise[0], noise[1]) * sigma
        out[c0:c1] = acc.to(torch.complex64)
    return out


def quantize_cs16(x: torch.Tensor) -> torch.Tensor:
    """complex -> interleaved int16 pairs (full scale 1.0 = 32768), as uint8 bytes."""
    iq = torch.view_as_real(x).reshape(-1) * 32768.0
    q = torch.clamp(torch.round(iq), -32768, 32767).to(torch.int16)
    return q.view(torch.uint8)


def quantize_cu8(x: torch.Tensor) -> torch.Tensor:
    """complex -> offset-binary uint8 pairs (code = 127.5 + 127.5 * x)."""
    iq = torch.view_as_real(x).reshape(-1) * 127.5 + 127.5
    return torch.clamp(torch.round(iq), 0, 255).to(torch.uint8)


# ---------------------------------------------------------------------------
# Device render (rvx_synth): the same scene, one thread per sample.
# ---------------------------------------------------------------------------

_KIND = {"cw": 0, "am": 1, "nfm": 2, "wfm": 2}


def carrier_table(carriers: list[Carrier]):
    """The scene as the C ABI's rvx_carrier array."""
    from . import _lib

    arr = (_lib.Carrier * max(1, len(carriers)))()
    for i, c in enumerate(carriers):
        index = (c.deviation_hz / c.tone_hz) if c.kind in ("nfm", "wfm") else 0.0
        arr[i] = _lib.Carrier(_KIND[c.kind], 0, int(c.freq_hz), int(c.tone_hz),
                              10.0 ** (c.level_dbfs / 20.0), index, float(c.depth))
    return arr


def render_device(dev, carriers: list[Carrier], rate: int, start: int, count: int,
                  noise_dbfs: float | None, seed: int, fmt: int, out: torch.Tensor) -> None:
    """Render samples start .. start+count into `out` (a CUDA tensor): native
    cs16 / cu8 codes, or cf32 with fmt = -1."""
    import ctypes as C

    from . import _lib

    table = carrier_table(carriers)
    sigma = 0.0 if noise_dbfs is None else 10.0 ** (noise_dbfs / 20.0) / math.sqrt(2.0)
    p = _lib.SynthParams(rate, start, count, seed & (2 ** 64 - 1), sigma, fmt, 0)
    _lib.check(dev.lib.rvx_synth(dev.ctx, table, len(carriers), C.byref(p), out.data_ptr(),
                                 out.numel() * out.element_size()))


def render_numpy(carriers: list[Carrier], rate: int, start: int, count: int,
                 noise_dbfs: float | None, seed: int) -> np.ndarray:
    """FP64 CPU evaluation of exactly what rvx_synth computes, before rounding
    to float (for spot checks of the device render)."""
    n = np.arange(start, start + count, dtype=np.uint64)
    acc = np.zeros(count, dtype=np.complex128)

    def turns(f):
        fr = np.uint64(int(f) % rate)
        return ((fr * (n % np.uint64(rate))) % np.uint64(rate)).astype(np.float64) / rate

    for c in carriers:
        amp = 10.0 ** (c.level_dbfs / 20.0)
        ph = turns(c.freq_hz)
        env = amp
        if c.kind in ("nfm", "wfm"):
            ph = ph + (c.deviation_hz / c.tone_hz) * np.sin(2 * np.pi * turns(c.tone_hz)) / (
                2 * np.pi)
        elif c.kind == "am":
            env = amp * (1.0 + c.depth * np.cos(2 * np.pi * turns(c.tone_hz))) / (1.0 + c.depth)
        acc += env * np.exp(2j * np.pi * ph)
    if noise_dbfs is not None:
        sigma = 10.0 ** (noise_dbfs / 20.0) / math.sqrt(2.0)

        def mix(x):
            x = x + np.uint64(0x9E3779B97F4A7C15)
            x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return x ^ (x >> np.uint64(31))

        with np.errstate(over="ignore"):
            h = mix(np.uint64(seed & (2 ** 64 - 1)) ^ mix(n))
            h2 = mix(h)
        u1 = ((h >> np.uint64(11)) + np.uint64(1)).astype(np.float64) / 9007199254740992.0
        u2 = ((h2 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) / 9007199254740992.0
        r = np.sqrt(-2.0 * np.log(u1))
        acc += sigma * r * np.exp(2j * np.pi * u2)
    return acc
