"""Seeded burst scenes for the learned detector (synthetic inputs, rendered on
the GPU with torch as plumbing).

A scene is 1-40 bursts across the 122.88 MS/s span: CW, 2-FSK, 4-FSK, an
OFDM-like multicarrier (64 subcarriers, QPSK, a new symbol every 64/bw s) and
a linear chirp; durations 1 ms - 2 s (log-uniform), bandwidths 2 kHz - 2 MHz
(log-uniform), SNR -5..30 dB in the burst's own bandwidth. Each scene renders
twice from the same draws: with complex white noise, and noise-free (the copy
the plain detector labels). Bursts key up and down with 100 us raised-cosine
ramps.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

KINDS = ("cw", "fsk2", "fsk4", "ofdm", "chirp")
RAMP_S = 100e-6


@dataclass
class Burst:
    kind: str
    freq_hz: float     # centre offset from the tuned centre
    bw_hz: float
    t0: float          # start, seconds from the scene's first sample
    dur: float
    snr_db: float
    seed: int


def draw_scene(seed: int, rate: float, span_s: float, max_bursts: int = 40) -> list[Burst]:
    g = np.random.default_rng(seed)
    out = []
    for i in range(int(g.integers(1, max_bursts + 1))):
        kind = KINDS[int(g.integers(0, len(KINDS)))]
        bw = float(np.exp(g.uniform(np.log(2e3), np.log(2e6))))
        if kind == "ofdm":
            bw = max(bw, 2e5)  # 64 subcarriers need room
        if kind == "cw":
            bw = 2e3
        f = float(g.uniform(-0.45 * rate + bw, 0.45 * rate - bw))
        dur = float(np.exp(g.uniform(np.log(1e-3), np.log(2.0))))
        t0 = float(g.uniform(-dur, span_s))
        out.append(Burst(kind, f, bw, t0, dur, float(g.uniform(-5.0, 30.0)),
                         int(g.integers(0, 2 ** 62))))
    return out


def _waveform(b: Burst, t: torch.Tensor, rate: float) -> torch.Tensor:
    """Unit-power complex baseband of burst b at times t (float64, seconds)."""
    g = torch.Generator(device=t.device)
    g.manual_seed(b.seed)
    tau = t - b.t0
    two_pi = 2.0 * math.pi
    if b.kind == "cw":
        ph = two_pi * b.freq_hz * tau
    elif b.kind in ("fsk2", "fsk4"):
        m = 2 if b.kind == "fsk2" else 4
        baud = b.bw_hz / 4.0
        nsym = int(math.ceil(float(tau[-1]) * baud)) + 2
        sym = torch.randint(0, m, (nsym,), generator=g, device=t.device).double()
        tones = (2.0 * sym - (m - 1)) / (m - 1) * (b.bw_hz / 2.0 - baud / 2.0)
        idx = torch.clamp((tau * baud).long(), 0, nsym - 1)
        inst = b.freq_hz + tones[idx]
        # continuous phase; the running sum on the host, in order, because a
        # device scan's addition order (and so its rounding) can vary run to run
        ph = (two_pi * torch.cumsum(inst.cpu(), 0) / rate).to(t.device)
    elif b.kind == "chirp":
        k = b.bw_hz / b.dur
        ph = two_pi * ((b.freq_hz - b.bw_hz / 2.0) * tau + 0.5 * k * tau * tau)
    else:  # ofdm-like: 64 QPSK subcarriers, symbols of 64/bw seconds
        spacing = b.bw_hz / 64.0
        nsym = int(math.ceil(float(tau[-1]) * spacing)) + 2
        idx = torch.clamp((tau * spacing).long(), 0, nsym - 1)
        x = torch.zeros(t.numel(), dtype=torch.complex128, device=t.device)
        for k in range(64):
            q = torch.randint(0, 4, (nsym,), generator=g, device=t.device).double()
            phase = two_pi * (b.freq_hz + (k - 31.5) * spacing) * tau + (math.pi / 2) * q[idx] \
                + math.pi / 4
            x += torch.polar(torch.ones_like(phase), phase)
        return (x / 8.0).to(torch.complex64)
    return torch.polar(torch.ones_like(ph), ph).to(torch.complex64)


def render(bursts: list[Burst], rate: float, n: int, noise_sigma: float, noisy: bool,
           seed: int, device="cuda") -> torch.Tensor:
    """Complex64 samples 0..n-1 of the scene; noise_sigma per component."""
    x = torch.zeros(n, dtype=torch.complex64, device=device)
    n0 = 2.0 * noise_sigma ** 2 / rate  # noise power per Hz
    for b in bursts:
        i0 = max(0, int(math.ceil(b.t0 * rate)))
        i1 = min(n, int(math.floor((b.t0 + b.dur) * rate)))
        if i1 <= i0:
            continue
        t = torch.arange(i0, i1, dtype=torch.float64, device=device) / rate
        amp = math.sqrt(10.0 ** (b.snr_db / 10.0) * n0 * b.bw_hz)
        # raised-cosine key-up and key-down of RAMP_S (a transmitter's, and it
        # keeps switching splatter out of the labels)
        ramp = min(RAMP_S, b.dur / 4.0)
        tau = t - b.t0
        edge = torch.clamp(torch.minimum(tau, b.dur - tau) / ramp, 0.0, 1.0)
        env = torch.sin(0.5 * math.pi * edge) ** 2
        x[i0:i1] += (amp * env).to(torch.float32) * _waveform(b, t, rate)
    if noisy:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        x += torch.complex(torch.randn(n, generator=g, device=device),
                           torch.randn(n, generator=g, device=device)) * noise_sigma
    return x


def to_cs16(x: torch.Tensor) -> torch.Tensor:
    """Complex samples (full scale 1) -> cs16 words (int16 I, Q interleaved) as uint8."""
    iq = torch.view_as_real(x).mul(32767.0).round().clamp(-32768, 32767).to(torch.int16)
    return iq.contiguous().view(torch.uint8).reshape(-1)
