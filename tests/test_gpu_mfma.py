"""GPU: the matrix-core channelizer's precision against FP64 (configuration 4).

A 100 dB dynamic-range multitone (one -3 dBFS tone, 50 tones at -100 dBFS,
AWGN) through the 4096-channel grid: the plain fp32 chain and the split-bf16
(hi + lo, three products) matrix-core form, each against the float64
channelizer. Reported per form: the worst error relative to the
carrier (dBc), the spur-free dynamic range of the error (carrier power over
the strongest error in any other channel) and per-channel SNR against FP64.
The display spectrum in FP8 e4m3 is checked against its definition.
"""
import json

import numpy as np
import pytest

from helpers import from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib, precision
from revenant_amd.dsp import FORMAT_CS16, GridParams, build_twiddles, design_prototype

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

M, L, D = precision.M, precision.L, precision.D
BLOCKS = precision.BLOCKS
multitone, metrics = precision.multitone, precision.metrics


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
    for name, prec in (("plain_fp32", None), ("bf16x3", _lib.PREC_BF16X3)):
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


def test_mfma_follows_the_exact_chain_through_a_wrapping_ring(dev):
    """Odd first block (chunks straddle the ring's 4-slot runs), a ring that
    wraps, 16 chunks over the XCDs: every output and frame partial within
    the split-bf16 form's error of the exact chain's."""
    grid = GridParams(M, L, D)
    taps = design_prototype(grid, beta=8.0).taps
    blocks, ring_blocks = 64, 64
    first = (M * L - 1 + D - 1) // D + 5
    n = (first + blocks) * D + 3
    codes = rng(0xC4C4C4C400000002).integers(-30000, 30000, 2 * n, dtype=np.int16)
    packed = to_dev(codes.view(np.uint8))
    tt, tw = to_dev(taps), to_dev(build_twiddles(M).view(np.float32))
    desc = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, blocks, ring_blocks, 4,
                          np.float32(1 / 16))
    rings, parts = [], []
    for prec in (None, _lib.PREC_BF16X3):
        ring = torch.zeros(2 * M * ring_blocks, dtype=torch.float32, device="cuda")
        part = torch.zeros((blocks // 4) * M, dtype=torch.float32, device="cuda")
        if prec is None:
            dev.chain(desc, packed, tt, tw, ring, part)
        else:
            dev.chain_mfma(desc, prec, packed, tt, tw, ring, part)
        rings.append(from_dev(ring).astype(np.float64))
        parts.append(from_dev(part).astype(np.float64))
    exact, fast = rings
    scale = np.max(np.abs(exact))
    print(f"mfma vs exact: max |diff| / max |exact| = {np.max(np.abs(fast - exact)) / scale:.3e}")
    assert np.max(np.abs(fast - exact)) <= 2e-5 * scale
    assert np.max(np.abs(parts[1] - parts[0])) <= 1e-4 * np.max(parts[0])


def test_mfma_refuses_the_withdrawn_fp16_form_and_other_grids(dev):
    grid = GridParams(M, L, D)
    taps = to_dev(design_prototype(grid, beta=8.0).taps)
    tw = to_dev(build_twiddles(M).view(np.float32))
    first = (M * L - 1 + D - 1) // D
    n = (first + 8) * D + 1
    packed = torch.zeros(4 * n, dtype=torch.uint8, device="cuda")
    ring = torch.zeros(2 * M * 8, dtype=torch.float32, device="cuda")
    part = torch.zeros(2 * M, dtype=torch.float32, device="cuda")
    desc = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, 8, 8, 4, np.float32(1 / 16))
    with pytest.raises(_lib.RvxError, match="fp16 form is withdrawn"):
        dev.chain_mfma(desc, _lib.PREC_FP16, packed, taps, tw, ring, part)
    odd = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, 6, 8, 0, np.float32(1 / 16))
    with pytest.raises(_lib.RvxError, match="chunks"):
        dev.chain_mfma(odd, _lib.PREC_BF16X3, packed, taps, tw, ring, part)


def test_mfma_ring_pitch_moves_rows_only(dev):
    """Rows ring_blocks + 32 apart hold exactly what the unpadded ring holds,
    and the padding is not written."""
    grid = GridParams(M, L, D)
    taps = design_prototype(grid, beta=8.0).taps
    blocks = ring_blocks = 16
    pitch = ring_blocks + _lib.RING_PITCH_PAD
    first = (M * L - 1 + D - 1) // D + 2
    n = (first + blocks) * D + 1
    codes = rng(0xC4C4C4C400000003).integers(-30000, 30000, 2 * n, dtype=np.int16)
    packed = to_dev(codes.view(np.uint8))
    tt, tw = to_dev(taps), to_dev(build_twiddles(M).view(np.float32))
    out = []
    for p in (0, pitch):
        desc = _lib.ChainDesc(FORMAT_CS16, M, L, D, 0, n, first, blocks, ring_blocks, 4,
                              np.float32(1 / 16))
        desc.ring_pitch = p
        stride = p or ring_blocks
        ring = torch.full((2 * M * stride,), 3.0, dtype=torch.float32, device="cuda")
        part = torch.zeros((blocks // 4) * M, dtype=torch.float32, device="cuda")
        dev.chain_mfma(desc, _lib.PREC_BF16X3, packed, tt, tw, ring, part)
        out.append(from_dev(ring).reshape(M, stride, 2))
    assert np.array_equal(out[1][:, :ring_blocks].view(np.uint32), out[0].view(np.uint32))
    assert np.all(out[1][:, ring_blocks:] == 3.0)
