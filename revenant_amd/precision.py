"""The precision study of the matrix-core channelizer (configuration 4).

A 100 dB dynamic-range multitone through the 4096-channel grid: one tone at
-3 dBFS centred 1234.5 Hz off channel 37, fifty tones at -100 dBFS at seeded
channels and offsets, complex AWGN at -120 dBFS, quantised to cs16. Each
form's channel outputs are compared with a float64 channelizer of the same
codes (supplied by the caller) by ``metrics``: the worst error relative to the
carrier (dBc), the spur-free dynamic range of the error (carrier power over
the largest mean error power of any other channel), per-channel SNR, and the
largest error relative to the largest output.
"""
from __future__ import annotations

import numpy as np

M, L, D = 4096, 16, 2048
BLOCKS = 8
RATE = 122_880_000
SEED = 0xC4C4C4C400000001
CARRIER_CHANNEL = 37


def multitone(seed: int = SEED, blocks: int = BLOCKS):
    """(cs16 words, first block with full history, samples, carrier channel,
    the weak tones' channels)."""
    g = np.random.default_rng(seed)
    first = (M * L - 1 + D - 1) // D
    n = (first + blocks) * D + 1
    t = np.arange(n)
    carrier = CARRIER_CHANNEL * RATE / M + 1234.5
    x = 10 ** (-3 / 20) * np.exp(2j * np.pi * carrier * t / RATE)
    tones = g.choice(np.arange(100, M - 100), 50, replace=False)
    for k in tones:
        f = (k if k < M // 2 else k - M) * RATE / M + float(g.uniform(-3000, 3000))
        x += 1e-5 * np.exp(2j * np.pi * f * t / RATE + 1j * g.uniform(0, 2 * np.pi))
    x += 10 ** (-120 / 20) * (g.standard_normal(n) + 1j * g.standard_normal(n)) / np.sqrt(2)
    iq = np.empty(2 * n)
    iq[0::2], iq[1::2] = x.real, x.imag
    codes = np.clip(np.round(iq * 32767), -32768, 32767).astype(np.int16)
    return codes.view(np.uint32), first, n, CARRIER_CHANNEL, np.sort(tones)


def metrics(got: np.ndarray, ref: np.ndarray, carrier: int, weak: np.ndarray) -> dict:
    """got, ref: complex [blocks, M]."""
    err = got - ref
    peak = np.max(np.abs(ref[:, carrier]))
    err_dbc = 20 * np.log10(np.max(np.abs(err)) / peak)
    perr = np.mean(np.abs(err) ** 2, axis=0)
    others = np.delete(perr, carrier)
    sfdr = 10 * np.log10(np.mean(np.abs(ref[:, carrier]) ** 2) / np.max(others))
    psig = np.mean(np.abs(ref) ** 2, axis=0)
    snr = 10 * np.log10(psig / np.maximum(perr, 1e-300))
    return {"err_dbc": round(float(err_dbc), 2), "sfdr_db": round(float(sfdr), 2),
            "snr_carrier_db": round(float(snr[carrier]), 2),
            "snr_weak_tones_min_db": round(float(snr[weak].min()), 2),
            "snr_weak_tones_median_db": round(float(np.median(snr[weak])), 2),
            "max_rel_error": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))}


def ring_outputs(ring: np.ndarray, first: int, blocks: int, ring_blocks: int) -> np.ndarray:
    """The channel ring (float32 pairs, M x ring_blocks) as complex [blocks, M]."""
    r = ring.reshape(M, ring_blocks, 2)
    slots = [(first + b) % ring_blocks for b in range(blocks)]
    return (r[:, slots, 0] + 1j * r[:, slots, 1]).T.astype(np.complex128)


def run_forms(dev, torch, ref: np.ndarray, forms) -> dict:
    """Each (name, precision) form of the 4096-channel chain over the multitone
    (precision None: the bit-exact fp32 chain), against ``ref`` (complex128
    [BLOCKS, M], the float64 channelizer of ``multitone()``'s codes)."""
    from . import _lib
    from .dsp import FORMAT_CS16, GridParams, build_twiddles, design_prototype

    codes, first, n, carrier, weak = multitone()
    taps = design_prototype(GridParams(M, L, D), beta=8.0).taps
    desc = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, BLOCKS, BLOCKS, 4,
                          np.float32(1 / 16))
    packed = torch.from_numpy(codes.view(np.uint8).copy()).cuda()
    tt = torch.from_numpy(taps).cuda()
    tw = torch.from_numpy(build_twiddles(M).view(np.float32)).cuda()
    out = {}
    for name, prec in forms:
        ring = torch.zeros(2 * M * BLOCKS, dtype=torch.float32, device="cuda")
        part = torch.zeros(2 * M, dtype=torch.float32, device="cuda")
        if prec is None:
            dev.chain(desc, packed, tt, tw, ring, part)
        else:
            dev.chain_mfma(desc, prec, packed, tt, tw, ring, part)
        torch.cuda.synchronize()
        got = ring_outputs(ring.cpu().numpy(), first, BLOCKS, BLOCKS)
        out[name] = metrics(got, ref, carrier, weak)
    return out
