"""GPU: impulse excision inside the streaming engine, between the converted
samples and the channelizer (core/engine/graph.cpp:1490,
record_impulse_excision).

The engine converts each push, runs the excision cursor's four passes in place
on its sample ring and channelizes the cleaned samples (the chain's cf32
input). The oracle runs the same stream through the excision twins block by
block (tests/reference/test_impulse_excision.cpp:113-196; the samples are
cs16 codes, whose conversion is exact), then the branch and transform twins
over the cleaned samples, the averaged-spectrum twin, and each receiver's fine
and demod twins. Everything must agree bit for bit: the channel ring the
engine holds, the averaged spectrum, the excision's counters and every
receiver's audio, however the stream is cut into pushes.
"""
import numpy as np
import pytest

from excision_stream import OraclePasses, impulsive_capture, run_stream
from helpers import assert_bit_exact, rng
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, FORMAT_CS16, GridParams, VrxRequest,
                              build_twiddles, design_prototype, ema_params, plan_vrx)
from revenant_amd.excision import design_excision, excision_threshold

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
GRID = GridParams(64, 16, 32)
SAMPLES = 1 << 19
FRAME, ROWS = 4, 2
ALPHA = np.float32(1.0 / 16.0)
THRESHOLD_DB = 15.0
RING_BLOCKS = 4096
REQUESTS = [VrxRequest(200_000, DEMOD_NFM, 12500), VrxRequest(-333_333, DEMOD_AM, 10000)]


@pytest.fixture(scope="module")
def stream():
    cfg = design_excision(RATE)
    x = impulsive_capture(SAMPLES, 0x4558454E47, 0.05, 0.0123, 300, cfg.max_width)
    g = rng(0x4558454E48)
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 0.08), (-333_333, 0.05)):
        x = x + (a * np.exp(2j * np.pi * f * t / RATE)).astype(np.complex64)
    codes = np.empty(2 * SAMPLES, dtype=np.int16)
    codes[0::2] = np.clip(np.round(x.real * 8000.0), -32768, 32767)
    codes[1::2] = np.clip(np.round(x.imag * 8000.0), -32768, 32767)
    del g
    converted = (codes.astype(np.float32) * np.float32(2.0 ** -15)).view(np.complex64)
    plans = [plan_vrx(GRID, RATE, r) for r in REQUESTS]
    return codes.view(np.uint32), converted, plans


@pytest.fixture(scope="module")
def oracle_stream(stream):
    """Excision twins over the whole stream, then branch + transform twins over
    the cleaned samples for every block the engine can run, and the averaged
    spectrum twin."""
    _, converted, _ = stream
    cleaned, _, counters, cleaned_end, _ = run_stream(converted, RATE, 65536, THRESHOLD_DB,
                                                      OraclePasses())
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
    return first, blocks, cap_b, chan.reshape(M, cap_b, 2), state, counters


def _run(dev, stream, cuts, ring_out=False):
    from revenant_amd.engine import Engine

    codes, _, plans = stream
    eng = Engine(dev, GRID, design_prototype(GRID, beta=8.0).taps, FORMAT_CS16, max_push=65536,
                 ring_blocks=RING_BLOCKS, frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA))
    eng.set_excision(design_excision(RATE), excision_threshold(THRESHOLD_DB))
    ids = [eng.add(p, audio_capacity=1 << 16) for p in plans]
    pos = 0
    for n in cuts:
        eng.push(codes[pos:pos + n])
        pos += n
    assert pos == codes.size
    audio = [eng.read_audio(i) for i in ids]
    st = eng.stats()
    ring = None
    if ring_out:
        ring = np.empty(2 * GRID.channels * RING_BLOCKS, dtype=np.float32)
        _lib.check(dev.lib.rvx_buf_read(dev.ctx, ring.ctypes.data, eng.channel_ring(),
                                        ring.nbytes))
        ring = ring.reshape(GRID.channels, RING_BLOCKS, 2)
    out = (audio, eng.spectrum(), eng.excision_counters(), (st.blocks, st.next_block), ring)
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


def test_engine_excision_matches_the_oracle(dev, stream, oracle_stream):
    from test_gpu_engine import _oracle_audio

    first, blocks, cap_b, chan, state, counters = oracle_stream
    audio, spec, got_counters, (nb, next_block), ring = _run(
        dev, stream, _cuts(0x45, SAMPLES, 3000, 65536), ring_out=True)
    assert nb == blocks and next_block == first + blocks
    assert got_counters[0] > 0 and got_counters[1] > 0, "impulses were excised"
    assert list(got_counters[:3]) == [int(c) for c in counters[:3]]
    # the ring's last RING_BLOCKS blocks against the twins'
    for m in range(next_block - RING_BLOCKS, next_block):
        assert_bit_exact(ring[:, m & (RING_BLOCKS - 1)], chan[:, m & (cap_b - 1)],
                         what=f"engine channel ring, block {m}")
    assert_bit_exact(spec, state, what="averaged spectrum after the stream")
    _, _, plans = stream
    for v, (p, (a0, got)) in enumerate(zip(plans, audio)):
        ref = _oracle_audio(p, chan, cap_b, a0, got.size // p.demod.channels)
        assert_bit_exact(got, ref, what=f"receiver {v} after excision vs oracle")
        assert np.abs(got).max() > 0


def test_engine_excision_is_blocking_invariant(dev, stream):
    a = _run(dev, stream, _cuts(0x46, SAMPLES, 1, 9000))
    b = _run(dev, stream, [65536] * (SAMPLES // 65536))
    assert a[3] == b[3]
    assert list(a[2]) == list(b[2])
    assert_bit_exact(a[1], b[1], what="spectrum, ragged vs even pushes")
    for v, ((a0, x), (b0, y)) in enumerate(zip(a[0], b[0])):
        assert a0 == b0
        n = min(x.size, y.size)
        assert n > x.size - 64
        assert_bit_exact(x[:n], y[:n], what=f"receiver {v}, ragged vs even pushes")


def test_engine_excision_is_set_before_the_first_push(dev, stream):
    from revenant_amd.engine import Engine

    codes = stream[0]
    eng = Engine(dev, GRID, design_prototype(GRID, beta=8.0).taps, FORMAT_CS16, max_push=65536,
                 ring_blocks=RING_BLOCKS)
    eng.push(codes[:1000])
    with pytest.raises(_lib.RvxError):
        eng.set_excision(design_excision(RATE), excision_threshold(THRESHOLD_DB))
    eng.close()
