"""GPU: a second grid in the streaming engine for receivers wider than a
channel of the engine's own (rvx_engine_add_grid, rvx_engine_add_vrx_on).

The reference runs one grid and coarsens it until the widest receiver fits
(core/engine/engine.h:160-200, engine::channel_count_for: 8 channels carry a
200 kHz receiver at 2.4 MS/s); the engine keeps its 64-channel grid for the
spectrum and the narrow receivers and adds the 8-channel one beside it. A
receiver on the added grid must equal the oracle's composition bit for bit:
the chain twin of the added grid over every block, then the receiver's fine
and demod twins (tests/reference/test_vrx.cpp:1241, blocking invariance; the
engine tests' oracle comparison), however the stream is cut, with rows taken
and given back as receivers come and go.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, rng
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, DEMOD_WFM, FORMAT_CS16, GridParams, VrxRequest,
                              build_nco_table, build_twiddles, channel_count_for,
                              design_prototype, ema_frame_decay, plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
MAIN = GridParams(64, 16, 32)
WIDE = GridParams(channel_count_for(RATE, 200_000), 16, channel_count_for(RATE, 200_000) // 2)
SAMPLES = 1 << 20
MAIN_RX = [VrxRequest(200_000, DEMOD_NFM, 12500), VrxRequest(-333_333, DEMOD_AM, 10000)]
WIDE_RX = [VrxRequest(-150_000, DEMOD_WFM, 200_000),   # channel 7, stereo
           VrxRequest(-160_000, DEMOD_NFM, 12500),     # channel 7 as well: shares its row
           VrxRequest(600_000, DEMOD_WFM, 200_000)]    # channel 2


@pytest.fixture(scope="module")
def stream():
    g = rng(0x4752494453000001)
    codes = g.integers(-6000, 6000, size=2 * SAMPLES, dtype=np.int16).astype(np.float64)
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 6000), (-333_333, 4000), (-160_000, 3000)):
        z = a * np.exp(2j * np.pi * f * t / RATE)
        codes[0::2] += z.real
        codes[1::2] += z.imag
    # two FM stations at 75 kHz deviation: a pilot-tone stereo multiplex
    # (1 kHz left, 3 kHz right) at -150 kHz and a 2 kHz mono tone at +600 kHz
    left, right = np.sin(2 * np.pi * 1000.0 * t / RATE), np.sin(2 * np.pi * 3000.0 * t / RATE)
    wp = 2 * np.pi * 19_000.0 * t / RATE
    for f, m in ((-150_000, 0.4 * ((left + right) / 2 + (left - right) / 2 * np.sin(2 * wp)) +
                  0.1 * np.sin(wp)),
                 (600_000, 0.8 * np.sin(2 * np.pi * 2000.0 * t / RATE))):
        phase = 2 * np.pi * 75_000.0 * np.cumsum(m) / RATE
        z = 5000 * np.exp(1j * (2 * np.pi * f * t / RATE + phase))
        codes[0::2] += z.real
        codes[1::2] += z.imag
    codes = np.clip(np.round(codes), -32768, 32767).astype(np.int16)
    main_plans = [plan_vrx(MAIN, RATE, r) for r in MAIN_RX]
    wide_plans = [plan_vrx(WIDE, RATE, r) for r in WIDE_RX]
    return codes.view(np.uint32), main_plans, wide_plans


def _engine(dev, max_rows=4, ring_blocks=1 << 15, pad=0):
    from revenant_amd.engine import Engine

    eng = Engine(dev, MAIN, design_prototype(MAIN, beta=8.0).taps, FORMAT_CS16, max_push=65536,
                 ring_blocks=4096, frame_blocks=4, row_frames=2,
                 ring_pitch=4096 + pad if pad else 0)
    gid = eng.add_grid(WIDE, design_prototype(WIDE, beta=8.0).taps, ring_blocks=ring_blocks,
                       max_rows=max_rows, ring_pitch=ring_blocks + pad if pad else 0)
    return eng, gid


def _cuts(seed, total, lo, hi):
    g = rng(seed)
    out, left = [], total
    while left:
        n = int(min(left, g.integers(lo, hi + 1)))
        out.append(n)
        left -= n
    return out


def _push(eng, codes, lo, hi, step=65536):
    for at in range(lo, hi, step):
        eng.push(codes[at:min(hi, at + step)])


@pytest.fixture(scope="module")
def wide_oracle(stream):
    """The added grid's chain twin over every block of the stream."""
    codes = stream[0]
    M, D, N = WIDE.channels, WIDE.decimation, WIDE.prototype_length
    first = (N - 1 + D - 1) // D
    blocks = (SAMPLES - 1) // D - first + 1
    cap = 1 << (blocks - 1).bit_length()
    od = oracle.OChain(FORMAT_CS16, M, WIDE.taps_per_branch, D, first, blocks, cap, 0,
                       np.float32(1.0 / 16.0), ema_frame_decay(np.float32(1.0 / 16.0), 4))
    ring, _ = oracle.chain(od, codes.view(np.uint8), design_prototype(WIDE, beta=8.0).taps,
                           build_twiddles(M), threads=4)
    return first, blocks, cap, ring.reshape(M, cap, 2)


def _oracle_audio(p, ring, cap, a0, n):
    """Receiver p's audio a0 .. a0+n-1 from the twin ring: the fine twin in
    dispatches the 32-bit resampler recurrence allows, then the demod twin."""
    import config1_scene as c1

    r, hist = p.demod.decimation, p.history
    fine_count = (n - 1) * r + hist + 1
    out_cap = 1 << (fine_count - 1).bit_length()
    chan = np.ascontiguousarray(ring[p.plan.placement.channel]).ravel()
    nco = build_nco_table(16)
    fine = np.zeros(2 * out_cap, dtype=np.float32)
    for fp in c1.fine_dispatches(p, 0, cap - 1, out_cap - 1, a0 * r - hist, fine_count):
        part = oracle.vrx_fine(p.fine, fp, chan, p.fine_taps, nco, out_cap)
        o = (fp.out_offset + np.arange(fp.count)) & (out_cap - 1)
        fine[2 * o], fine[2 * o + 1] = part[2 * o], part[2 * o + 1]
    dp = p.demod_block(out_cap - 1, a0, n)
    return oracle.vrx_demod(p.demod, dp, fine, p.weights)


def _audio_vs_oracle(p, ring, cap, a0, got, what):
    ref = _oracle_audio(p, ring, cap, a0, got.size // p.demod.channels)
    assert_bit_exact(got, ref, what=what)
    assert np.abs(got).max() > 0


def test_wide_receivers_on_the_added_grid_match_the_oracle(dev, stream, wide_oracle):
    codes, main_plans, wide_plans = stream
    first, blocks, cap, ring = wide_oracle
    eng, gid = _engine(dev)
    ids = [eng.add(p, audio_capacity=1 << 16, grid=gid) for p in wide_plans]
    main_ids = [eng.add(p, audio_capacity=1 << 16) for p in main_plans]
    # the two receivers on channel 7 share one row; channel 2 has its own
    rows = {p.plan.placement.channel: eng.grid_row(gid, p.plan.placement.channel)
            for p in wide_plans}
    assert len(rows) == 2 and rows[7][0] != rows[2][0] and rows[7][1] == first
    pos = 0
    for n in _cuts(0x47, SAMPLES, 3000, 65536):
        eng.push(codes[pos:pos + n])
        pos += n
    st = eng.stats()
    assert st.grid_blocks == blocks and st.audio_lost == 0
    for v, (p, rid) in enumerate(zip(wide_plans, ids)):
        a0, got = eng.read_audio(rid)
        _audio_vs_oracle(p, ring, cap, a0, got, f"wide receiver {v} vs oracle")
    # the stereo station's channels differ (its pilot gate open)
    a0, got = eng.read_audio(ids[0])
    assert wide_plans[0].demod.channels == 2
    for rid in main_ids:  # the engine's own grid still runs its receivers
        a0, got = eng.read_audio(rid)
        assert got.size > 10000 and np.abs(got).max() > 0
    eng.close()


def test_added_grid_is_blocking_invariant(dev, stream):
    codes, _, wide_plans = stream
    outs = []
    for cuts, on_device in ((_cuts(0x48, SAMPLES, 1, 9000), False),
                            ([65536] * (SAMPLES // 65536), True)):
        eng, gid = _engine(dev)
        ids = [eng.add(p, audio_capacity=1 << 16, grid=gid) for p in wide_plans]
        src = torch.from_numpy(codes.view(np.int32)).cuda() if on_device else codes
        pos = 0
        for n in cuts:
            eng.push(src[pos:pos + n])
            pos += n
        outs.append([eng.read_audio(i) for i in ids])
        eng.close()
    for v, ((a0, x), (b0, y)) in enumerate(zip(*outs)):
        assert a0 == b0
        n = min(x.size, y.size)
        assert n > x.size - 64
        assert_bit_exact(y[:n], x[:n], what=f"wide receiver {v}, ragged vs even pushes")


def test_rows_are_taken_and_given_back(dev, stream, wide_oracle):
    """A receiver added mid-stream takes a cleared row whose stream starts at
    the grid's next block; its audio starts where its history lies inside
    that stream and equals the oracle's. A row stays while a receiver uses
    it, and a grid with every row taken refuses another channel."""
    codes, _, wide_plans = stream
    first, blocks, cap, ring = wide_oracle
    eng, gid = _engine(dev, max_rows=2)
    early = eng.add(wide_plans[1], audio_capacity=1 << 16, grid=gid)  # channel 7
    half = SAMPLES // 2
    _push(eng, codes, 0, half)
    late = eng.add(wide_plans[2], audio_capacity=1 << 16, grid=gid)   # channel 2, mid-stream
    row2, from2 = eng.grid_row(gid, 2)
    assert from2 > first
    with pytest.raises(_lib.RvxError):  # both rows taken: channel 0 has none
        eng.add(plan_vrx(WIDE, RATE, VrxRequest(10_000, DEMOD_NFM, 12500)), audio_capacity=4096,
                grid=gid)
    shared = eng.add(wide_plans[0], audio_capacity=1 << 16, grid=gid)  # channel 7's row
    eng.remove(early)  # channel 7 stays: the stereo receiver still uses it
    assert eng.grid_row(gid, 7)[1] == first
    _push(eng, codes, half, SAMPLES)
    # the oracle's ring holds the whole stream, so audio whose history reached
    # below the row's first block (cleared slots in the engine) would differ
    a0, got = eng.read_audio(late)
    _audio_vs_oracle(wide_plans[2], ring, cap, a0, got, "receiver added mid-stream")
    a0, got = eng.read_audio(shared)
    _audio_vs_oracle(wide_plans[0], ring, cap, a0, got, "receiver joining a held row")
    eng.remove(late)  # channel 2's row is free again: channel 0 takes it
    eng.add(plan_vrx(WIDE, RATE, VrxRequest(10_000, DEMOD_NFM, 12500)), audio_capacity=4096,
            grid=gid)
    eng.close()


def test_add_grid_refuses_what_the_window_cannot_hold(dev):
    from revenant_amd.engine import Engine

    eng = Engine(dev, MAIN, design_prototype(MAIN, beta=8.0).taps, FORMAT_CS16, max_push=65536,
                 ring_blocks=4096)
    big = GridParams(4096, 16, 2048)
    with pytest.raises(_lib.RvxError):
        eng.add_grid(big, np.zeros(big.prototype_length, np.float32), ring_blocks=64, max_rows=4)
    with pytest.raises(_lib.RvxError):  # max_rows beyond the channel count
        eng.add_grid(WIDE, design_prototype(WIDE, beta=8.0).taps, ring_blocks=1 << 15, max_rows=9)
    eng.close()


def test_padded_ring_rows_change_no_output(dev, stream, wide_oracle):
    """Both grids' rows ring_blocks + 32 apart: the added grid's receivers
    still equal the oracle, and the engine's own receivers and channel ring
    equal an unpadded engine's bit for bit."""
    import ctypes as C

    codes, main_plans, wide_plans = stream
    first, blocks, cap, ring = wide_oracle
    runs = []
    for pad in (0, _lib.RING_PITCH_PAD):
        eng, gid = _engine(dev, pad=pad)
        assert eng.ring_pitch == 4096 + pad
        ids = [eng.add(p, audio_capacity=1 << 16, grid=gid) for p in wide_plans]
        main_ids = [eng.add(p, audio_capacity=1 << 16) for p in main_plans]
        for at in range(0, SAMPLES, 50_000):
            eng.push(codes[at:at + 50_000])
        wide_audio = [eng.read_audio(i) for i in ids]
        main_audio = [eng.read_audio(i) for i in main_ids]
        pitch = eng.ring_pitch
        host = np.zeros(2 * MAIN.channels * pitch, dtype=np.float32)
        _lib.check(dev.lib.rvx_buf_read(dev.ctx, host.ctypes.data,
                                        C.c_void_p(eng.channel_ring()), host.nbytes))
        runs.append((wide_audio, main_audio, host.reshape(MAIN.channels, pitch, 2)[:, :4096]))
        eng.close()
    for v, (p, (a0, got)) in enumerate(zip(wide_plans, runs[1][0])):
        _audio_vs_oracle(p, ring, cap, a0, got, f"padded wide receiver {v} vs oracle")
    for (a0, x), (b0, y) in zip(runs[0][1], runs[1][1]):
        assert a0 == b0
        assert_bit_exact(y, x, what="main receiver, padded vs unpadded ring")
    assert_bit_exact(runs[1][2], runs[0][2], what="channel ring rows, padded vs unpadded")
