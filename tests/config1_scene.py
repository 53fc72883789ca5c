"""The plain-chain scene at 2.4 MS/s and its float64 receiver reference.

Input: 2^20 complex uint8 offset-binary samples at 2.4 MS/s, generated from a
seed: one WFM carrier at +200 kHz (75 kHz deviation, a 1 kHz tone) over
complex AWGN 20 dB below it. Channelizer: 64 channels critically sampled
(D = M), 16 taps per branch, Kaiser beta 8. The 200 kHz receiver cannot sit on
a 37.5 kHz channel (core/engine/vrx_place.cpp refuses it), so it is placed on
the grid channel_count_for picks for it at this rate (vrx_place.cpp:187): 8
channels, two times oversampled (D = 4, 600 kS/s channels), same prototype
shape, from the same codes.

``fp64_receiver`` evaluates the receiver's mix, rational resampling,
discriminator and audio filter in float64 from the float64 channelizer: the
same plan tables (fine taps interpolated between phases, audio weights, gain),
exact phase rotation exp(-j 2 pi phi / 2^64) instead of the 2^16-entry table,
libm atan2 instead of the det_* series. Its difference from the float32 chain
is what the receiver's audio error against float64 means.
"""
from __future__ import annotations

import numpy as np

from revenant_amd import dsp

RATE = 2_400_000
SAMPLES = 1 << 20
SEED = 0x434F4E4631000001
WFM_HZ = 200_000
PLAIN = dsp.GridParams(64, 16, 64)
BETA = 8.0


def scene_codes() -> np.ndarray:
    """cu8 pairs (uint8, 2 * SAMPLES)."""
    import torch

    from revenant_amd.synth import Carrier, quantize_cu8, render

    carrier = Carrier("wfm", WFM_HZ, -6.0, 1000, 75_000)
    x = render([carrier], RATE, 0, SAMPLES, -26.0, SEED, device="cpu")
    return quantize_cu8(x).numpy()


def receiver_grid() -> dsp.GridParams:
    m = dsp.channel_count_for(RATE, WFM_HZ)
    return dsp.GridParams(m, 16, m // 2)


def cu8_to_complex(codes: np.ndarray) -> np.ndarray:
    c = (codes.astype(np.float64) - 127.5) / 127.5
    return c[0::2] + 1j * c[1::2]


def first_block(grid: dsp.GridParams) -> int:
    return (grid.prototype_length - 1 + grid.decimation - 1) // grid.decimation


def fp64_receiver(plan, channel: np.ndarray, chan_first: int, first_output: int, count: int,
                  first_audio: int, audio_count: int) -> np.ndarray:
    """Mono audio in float64 from one channel's float64 samples
    (``channel[j]`` is channel sample chan_first + j)."""
    p = plan.plan
    cfg = plan.fine
    T, P = cfg.taps, cfg.phases
    taps = plan.fine_taps.astype(np.complex128)
    cr, dr = int(p.channel_rate), int(p.demod_rate)
    # the fine stage's index arithmetic, from fine_block's own params: output
    # j takes channel sample n_j (newest) and the phase frac_j / out_rate
    fp, n0 = plan.fine_block(0, (1 << 30) - 1, (1 << 30) - 1, first_output, 1)
    step_whole, step_rem, out_rate = int(fp.step_whole), int(fp.step_rem), int(fp.out_rate)
    i = np.arange(count, dtype=np.int64)
    acc = int(fp.frac0) + i * step_rem
    n = int(n0) + i * step_whole + acc // out_rate
    frac = acc % out_rate
    scaled = frac * P
    phase = scaled // out_rate
    blend = (scaled % out_rate) / out_rate
    k = np.arange(T)
    slot = phase[:, None] + k[None, :] * P
    tap = taps[slot] * (1.0 - blend[:, None]) + taps[slot + 1] * blend[:, None]
    x = channel[(n[:, None] - k[None, :]) - chan_first]
    filtered = (x * tap).sum(axis=1)
    delta = int(p.fine_nco_delta)
    j = first_output + i
    turns = np.array([((int(v) * delta) % (1 << 64)) / 2.0 ** 64 for v in j])
    fine = filtered * np.exp(-2j * np.pi * turns)  # the table holds exp(-j theta)
    del cr, dr
    # discriminator, then the decimating audio filter (ascending taps) and gain
    d = plan.demod
    w = plan.weights[:d.audio_taps].astype(np.float64)
    prod = fine[1:] * np.conj(fine[:-1])
    det = np.angle(prod)  # det[j - 1] belongs to fine index first_output + j
    R = d.decimation
    out = np.empty(audio_count)
    gain = float(plan.demod_block((1 << 20) - 1, first_audio, 1).gain)
    for a in range(audio_count):
        idx = (first_audio + a) * R - first_output  # position in `fine`
        seg = det[idx - 1 - np.arange(d.audio_taps)]
        out[a] = (seg * w).sum() * gain
    return out


def chain_desc(grid: dsp.GridParams, samples: int = SAMPLES, frame: int = 16):
    """(oracle OChain, first block, block count, ring blocks) for every block of
    the input with full history."""
    from oracle import oracle

    first = first_block(grid)
    blocks = (samples - 1) // grid.decimation + 1 - first
    blocks -= blocks % frame if frame else 0
    ring_blocks = 1 << (blocks - 1).bit_length()
    alpha = np.float32(1 / 16)
    od = oracle.OChain(dsp.FORMAT_CU8, grid.channels, grid.taps_per_branch, grid.decimation,
                       first, blocks, ring_blocks, frame, alpha,
                       dsp.ema_frame_decay(alpha, frame) if frame else 0.0)
    return od, first, blocks, ring_blocks


def receiver_spans(plan, first: int, blocks: int):
    """(first fine output, fine count, first audio, audio count) that keep
    every read inside the channel samples first .. first + blocks - 1 and the
    fine outputs computed."""
    T = plan.fine.taps
    cr, dr = int(plan.plan.channel_rate), int(plan.plan.demod_rate)
    f0 = -(-(first + T) * dr // cr) + 1
    last_in = first + blocks - 1
    count = (last_in - 1) * dr // cr - f0 - 1
    hist = plan.history
    R = plan.demod.decimation
    a0 = -(-(f0 + hist + 1) // R)
    a_count = (f0 + count - 1) // R - a0
    return f0, count, a0, a_count


FINE_DISPATCH = 8192  # outputs per fine dispatch (the recurrence is 32-bit)


def fine_dispatches(plan, chan_base: int, chan_mask: int, fine_mask: int, f0: int, count: int):
    """The fine stage's dispatch params over f0 .. f0 + count - 1, in order."""
    out = []
    for c in range(0, count, FINE_DISPATCH):
        n = min(FINE_DISPATCH, count - c)
        fp, _ = plan.fine_block(chan_base, chan_mask, fine_mask, f0 + c, n)
        out.append(fp)
    return out
