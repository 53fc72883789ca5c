"""GPU: comb cancellation inside the streaming engine (core/engine/graph.cpp:
1715, record_comb_cancel), alone and ahead of impulse excision, the
reference's record order.

The lines are given as the caller's comb tracker would hand them over: each
line's 0.32 phase at a reference sample, its per-sample step and complex
amplitude. The oracle subtracts them from the whole converted stream in one
dispatch of the subtract twin (comb_kernels_reference.cpp:143), runs the
excision twins over the result when excision is on, then the branch and
transform twins, the averaged-spectrum twin and the receivers' twins. The
engine's channel ring, averaged spectrum and audio must agree bit for bit,
however the stream is cut (the lines' phases advance by the absolute index),
and its projections of the last push must equal the project twin's over
those raw samples.
"""
import ctypes as C

import numpy as np
import pytest

from excision_stream import OraclePasses, impulsive_capture, run_stream
from helpers import assert_bit_exact, rng
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, FORMAT_CS16, GridParams, VrxRequest,
                              build_twiddles, comb_step, comb_table, design_prototype, ema_params,
                              plan_vrx)
from revenant_amd.excision import design_excision, excision_threshold

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
GRID = GridParams(64, 16, 32)
SAMPLES = 1 << 19
FRAME, ROWS = 4, 2
ALPHA = np.float32(1.0 / 16.0)
RING_BLOCKS = 4096
REQUESTS = [VrxRequest(200_000, DEMOD_NFM, 12500), VrxRequest(-333_333, DEMOD_AM, 10000)]
COMB = [(150_000.0, 0.06, 0.3), (300_000.0, 0.04, 1.7), (450_000.0, 0.03, -2.2)]  # f, |a|, phase
SCALE = 8000.0  # codes per unit of the synthesised signal


def _lines(phase_index: int):
    """The comb's lines with their phases at absolute sample phase_index."""
    arr = (_lib.CombLine * len(COMB))()
    for k, (f, amp, ph) in enumerate(COMB):
        step = comb_step(f, RATE)
        p0 = int(round(ph / (2 * np.pi) * 2.0 ** 32)) % (1 << 32)
        arr[k].phase0 = (p0 + step * phase_index) % (1 << 32)
        arr[k].step = step
        arr[k].amplitude[0] = np.float32(amp * SCALE / 32768.0)  # as the converted codes carry it
        arr[k].amplitude[1] = np.float32(0.0)
    return arr


@pytest.fixture(scope="module")
def stream():
    bare = impulsive_capture(SAMPLES, 0x434F4D42, 0.05, 0.0123, 200,
                             design_excision(RATE).max_width)
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 0.08), (-333_333, 0.05)):
        bare = bare + (a * np.exp(2j * np.pi * f * t / RATE)).astype(np.complex64)
    x = bare.copy()
    for f, amp, ph in COMB:
        x = x + (amp * np.exp(1j * (2 * np.pi * f * t / RATE + ph))).astype(np.complex64)

    def quantize(z):
        codes = np.empty(2 * SAMPLES, dtype=np.int16)
        codes[0::2] = np.clip(np.round(z.real * SCALE), -32768, 32767)
        codes[1::2] = np.clip(np.round(z.imag * SCALE), -32768, 32767)
        return codes, (codes.astype(np.float32) * np.float32(2.0 ** -15)).view(np.complex64)

    codes, converted = quantize(x)
    _, without_comb = quantize(bare)
    plans = [plan_vrx(GRID, RATE, r) for r in REQUESTS]
    return codes.view(np.uint32), converted, plans, without_comb


def _composition(cleaned, cleaned_end):
    M, D, N = GRID.channels, GRID.decimation, GRID.prototype_length
    first = (N - 1 + D - 1) // D
    blocks = (cleaned_end - 1) // D - first + 1
    blocks -= blocks % (FRAME * ROWS)
    cap = 1 << (SAMPLES - 1).bit_length()
    ring_in = np.zeros(2 * cap, dtype=np.float32)
    ring_in[:2 * SAMPLES] = cleaned.view(np.float32)
    taps = design_prototype(GRID, beta=8.0).taps
    branch = oracle.pfb_branch(GRID.c(), taps, ring_in, _lib.PfbBranchParams(cap - 1, first * D,
                                                                             blocks))
    cap_b = 1 << (first + blocks - 1).bit_length()
    chan = np.zeros(2 * M * cap_b, dtype=np.float32)
    oracle.pfb_fft(_lib.PfbFftParams(M, D, M.bit_length() - 1, first, blocks, cap_b, cap_b - 1),
                   branch, build_twiddles(M), chan)
    ep = ema_params(M, cap_b, first, FRAME, blocks // FRAME, ALPHA, ROWS)
    partial = oracle.ema_partial(ep, chan)
    state = np.zeros(M, dtype=np.float32)
    oracle.ema_combine(ep, partial, state)
    return first, blocks, cap_b, chan.reshape(M, cap_b, 2), state


def _oracle_subtract(converted):
    cap = 1 << (SAMPLES - 1).bit_length()
    p = _lib.CombParams(cap - 1, 0, SAMPLES, len(COMB), 1, cap - 1)
    cleaned, _ = oracle.comb_subtract(p, converted, _lines(0), comb_table(), cap)
    return cleaned.view(np.complex64)[:SAMPLES].copy()


def _run(dev, stream, cuts, excise):
    from revenant_amd.engine import Engine

    codes, _, plans, _ = stream
    eng = Engine(dev, GRID, design_prototype(GRID, beta=8.0).taps, FORMAT_CS16, max_push=65536,
                 ring_blocks=RING_BLOCKS, frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA))
    eng.set_comb(_lines(1000), phase_index=1000)  # phases stated at any sample
    if excise:
        eng.set_excision(design_excision(RATE), excision_threshold(15.0))
    ids = [eng.add(p, audio_capacity=1 << 16) for p in plans]
    pos = 0
    for n in cuts:
        eng.push(codes[pos:pos + n])
        pos += n
    audio = [eng.read_audio(i) for i in ids]
    st = eng.stats()
    ring = np.empty(2 * GRID.channels * RING_BLOCKS, dtype=np.float32)
    _lib.check(dev.lib.rvx_buf_read(dev.ctx, ring.ctypes.data, eng.channel_ring(), ring.nbytes))
    proj = eng.comb_projections()
    out = (audio, eng.spectrum(), (st.blocks, st.next_block),
           ring.reshape(GRID.channels, RING_BLOCKS, 2), proj)
    eng.close()
    return out


def _cuts(seed, total, lo, hi):
    g = rng(seed)
    out, left = [], total
    while left:
        n = int(min(left, g.integers(lo, hi + 1)))
        out.append(n)
        left -= n
    return out


def _check(stream, got, want, what):
    from test_gpu_engine import _oracle_audio

    first, blocks, cap_b, chan, state = want
    audio, spec, (nb, next_block), ring, _ = got
    assert nb == blocks and next_block == first + blocks, what
    for m in range(next_block - RING_BLOCKS, next_block):
        assert_bit_exact(ring[:, m & (RING_BLOCKS - 1)], chan[:, m & (cap_b - 1)],
                         what=f"{what}: channel ring, block {m}")
    assert_bit_exact(spec, state, what=f"{what}: averaged spectrum")
    for v, (p, (a0, x)) in enumerate(zip(stream[2], audio)):
        ref = _oracle_audio(p, chan, cap_b, a0, x.size // p.demod.channels)
        assert_bit_exact(x, ref, what=f"{what}: receiver {v}")


def test_engine_comb_matches_the_oracle(dev, stream):
    converted = stream[1]
    want = _composition(_oracle_subtract(converted), SAMPLES)
    cuts = _cuts(0x43, SAMPLES - 8192, 2000, 65536) + [8192]  # the last push in 16 sub-blocks
    got = _run(dev, stream, cuts, excise=False)
    _check(stream, got, want, "comb")
    # the last push's projections: the project twin over its raw samples
    proj, start, count = got[4]
    assert (start, count) == (SAMPLES - cuts[-1], cuts[-1])
    subs = min(count & -count, 16)
    assert proj.shape == (len(COMB), subs)
    cap = 1 << (SAMPLES - 1).bit_length()
    p = _lib.CombParams(cap - 1, start & (cap - 1), count, len(COMB), subs, cap - 1)
    raw = np.zeros(cap, dtype=np.complex64)
    raw[:SAMPLES] = converted
    want_proj = oracle.comb_project(p, raw, _lines(start), comb_table())
    assert_bit_exact(proj.ravel(), want_proj, what="comb projections of the last push")
    # the spurs are gone: each comb channel's averaged power comes back to
    # within 0.5 dB of the same stream synthesised without the comb
    plain = _composition(converted, SAMPLES)[4]
    bare = _composition(stream[3], SAMPLES)[4]
    for f, _, _ in COMB:
        k = int(round(f / RATE * GRID.channels)) % GRID.channels
        assert 10 * np.log10(plain[k] / bare[k]) > 5.0, f"the line at {f} Hz stands out"
        assert abs(10 * np.log10(got[1][k] / bare[k])) < 0.5, f"comb line at {f} Hz"



def test_engine_comb_then_excision_match_the_oracle(dev, stream):
    subtracted = _oracle_subtract(stream[1])
    cleaned, _, _, cleaned_end, _ = run_stream(subtracted, RATE, 65536, 15.0, OraclePasses())
    want = _composition(cleaned, cleaned_end)
    _check(stream, _run(dev, stream, _cuts(0x44, SAMPLES, 1, 20000), excise=True), want,
           "comb + excision")
