"""GPU: the matrix-core channelizer's precision against FP64 (configuration 4).

A 100 dB dynamic-range multitone (one -3 dBFS tone, 50 tones at -100 dBFS,
AWGN) through the 4096-channel grid: the plain fp32 chain, the split-bf16
(hi + lo, three products) and the fp16 matrix-core forms, each against the
float64 channelizer. Reported per form: the worst error relative to the
carrier (dBc), the spur-free dynamic range of the error (carrier power over
the strongest error in any other channel) and per-channel SNR against FP64.
The display spectrum in FP8 e4m3 is checked against its definition.
"""
import json

import numpy as np
import pytest

from helpers import from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import FORMAT_CS16, GridParams, build_twiddles, design_prototype

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

M, L, D = 4096, 16, 2048
BLOCKS = 8
RATE = 122_880_000


def multitone(seed):
    g = rng(seed)
    first = (M * L - 1 + D - 1) // D
    n = (first + BLOCKS) * D + 1
    t = np.arange(n)
    carrier = 37 * RATE / M + 1234.5
    x = 10 ** (-3 / 20) * np.exp(2j * np.pi * carrier * t / RATE)
    tones = g.choice(np.arange(100, M - 100), 50, replace=False)
    for k in tones:
        f = (k if k < M // 2 else k - M) * RATE / M + float(g.uniform(-3000, 3000))
        x += 1e-5 * np.exp(2j * np.pi * f * t / RATE + 1j * g.uniform(0, 2 * np.pi))
    x += 10 ** (-120 / 20) * (g.standard_normal(n) + 1j * g.standard_normal(n)) / np.sqrt(2)
    iq = np.empty(2 * n)
    iq[0::2], iq[1::2] = x.real, x.imag
    codes = np.clip(np.round(iq * 32767), -32768, 32767).astype(np.int16)
    return codes.view(np.uint32), first, n, 37, np.sort(tones)


def metrics(got, ref, carrier, weak):
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


def test_mfma_precision_against_fp64(dev):
    codes, first, n, carrier, weak = multitone(0xC4C4C4C400000001)
    grid = GridParams(M, L, D)
    taps = design_prototype(grid, beta=8.0).taps
    ref = oracle.fp64_chain(codes, taps, M, D, first, BLOCKS)
    desc = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, BLOCKS, BLOCKS, 4,
                          np.float32(1 / 16))
    packed, tt = to_dev(codes.view(np.uint8)), to_dev(taps)
    tw = to_dev(build_twiddles(M).view(np.float32))
    out = {}
    rings = {}
    for name, prec in (("plain_fp32", None), ("bf16x3", _lib.PREC_BF16X3),
                       ("fp16", _lib.PREC_FP16)):
        ring = torch.zeros(2 * M * BLOCKS, dtype=torch.float32, device="cuda")
        part = torch.zeros(2 * M, dtype=torch.float32, device="cuda")
        if prec is None:
            dev.chain(desc, packed, tt, tw, ring, part)
        else:
            dev.chain_mfma(desc, prec, packed, tt, tw, ring, part)
        r = from_dev(ring).reshape(M, BLOCKS, 2)
        slots = [(first + b) % BLOCKS for b in range(BLOCKS)]
        got = (r[:, slots, 0] + 1j * r[:, slots, 1]).T.astype(np.complex128)  # [blocks, M]
        out[name] = metrics(got, ref, carrier, weak)
        rings[name] = (got, from_dev(part))
    print("PRECISION " + json.dumps(out))
    assert out["plain_fp32"]["err_dbc"] < -120
    assert out["bf16x3"]["err_dbc"] <= -80, out
    assert out["fp16"]["err_dbc"] < -40, out
    # the averaged spectrum follows the channel values
    p64 = rings["plain_fp32"][1]
    pb = rings["bf16x3"][1]
    assert np.max(np.abs(pb - p64)) <= 1e-5 * np.max(p64)
    # FP8 display rows: e4m3 of dB / 4, within half a step of the format
    power = torch.from_numpy(p64).cuda()
    fp8 = torch.empty(power.numel(), dtype=torch.uint8, device="cuda")
    dev.spectrum_fp8(power, fp8)
    dec = oracle.e4m3_decode(from_dev(fp8)) * 4.0
    db = 10 * np.log10(np.maximum(p64.astype(np.float64), 1e-20))
    step = np.where(np.abs(db / 4) >= 2.0 ** -6, 2.0 ** (np.floor(np.log2(np.abs(db / 4))) - 3),
                    2.0 ** -9) * 4.0
    assert np.all(np.abs(dec - db) <= 0.5 * step + 1e-4 * np.abs(db)), np.max(np.abs(dec - db))
