"""GPU: a capture written to disk and played back through capture:// feeds
the engine exactly as the live stream did (after
tests/capture/test_writer.cpp:519-665, a written segment replays sample for
sample): averaged spectrum, block counts and every receiver's audio are
bit-identical to pushing the same codes directly, across segment boundaries
the writer's size limit put in."""
import numpy as np
import pytest

from helpers import rng
from revenant_amd.capture import CaptureReplay, CaptureWriter, SourceBlock, catalog, feed
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, DEMOD_USB, FORMAT_CS16, GridParams, VrxRequest,
                              design_prototype, plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
GRID = GridParams(64, 16, 32)
SAMPLES = 1 << 19


def _stream():
    g = rng(0x43415054)
    codes = g.integers(-9000, 9000, size=2 * SAMPLES, dtype=np.int16)
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 6000), (-333_333, 4000), (510_400, 3000)):
        z = a * np.exp(2j * np.pi * f * t / RATE)
        codes[0::2] = np.clip(codes[0::2] + z.real, -32768, 32767).astype(np.int16)
        codes[1::2] = np.clip(codes[1::2] + z.imag, -32768, 32767).astype(np.int16)
    return codes


def _engine(dev):
    from revenant_amd.engine import Engine
    proto = design_prototype(GRID, beta=8.0).taps
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096, frame_blocks=4,
                 row_frames=2, alpha=1.0 / 16.0)
    reqs = [VrxRequest(200_000, DEMOD_NFM, 12500), VrxRequest(-333_333, DEMOD_AM, 10000),
            VrxRequest(510_000, DEMOD_USB, 2700)]
    ids = [eng.add(plan_vrx(GRID, RATE, r), audio_capacity=1 << 16) for r in reqs]
    return eng, ids


def _outputs(eng, ids):
    audio = [eng.read_audio(i) for i in ids]
    st = eng.stats()
    return audio, eng.spectrum(), (st.blocks, st.next_block, st.spectrum_rows)


def test_a_capture_replays_into_the_engine_as_the_live_stream(dev, tmp_path):
    codes = _stream()
    eng, ids = _engine(dev)
    for at in range(0, SAMPLES, 65536):
        eng.push(codes[2 * at:2 * (at + 65536)].view(np.uint32))
    live = _outputs(eng, ids)
    eng.close()

    w = CaptureWriter(str(tmp_path), max_segment_bytes=600_000, stream_epoch=77)
    for at in range(0, SAMPLES, 40_000):
        n = min(40_000, SAMPLES - at)
        w.write(SourceBlock(codes[2 * at:2 * (at + n)], at, n, 100_000_000), RATE, "cs16",
                1_790_000_000_000_000_000)
    w.close("stop")
    assert len(catalog(tmp_path)) >= 3

    eng, ids = _engine(dev)
    print("capture replay into the engine", flush=True)
    pushed = feed(eng, CaptureReplay(str(tmp_path)).blocks(30_000), max_push=65536)
    assert pushed == SAMPLES
    replayed = _outputs(eng, ids)
    eng.close()

    for (f0, a0), (f1, a1) in zip(live[0], replayed[0]):
        assert f0 == f1 and a0.size > 1000
        assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
    assert np.array_equal(live[1].view(np.uint32), replayed[1].view(np.uint32))
    assert live[2] == replayed[2]


def test_spectrum_frames_and_audio_chunks_from_the_engine(dev):
    """The delivery contract over the engine: power_db is the spectrum twin
    of the engine's own channel ring, bit for bit; the percentiles are the
    levels twin of that frame; the geometry puts a +200 kHz tone's peak in
    the bin whose centre is nearest +200 kHz; start / count are source
    indices with the group delay off; sequences count frames delivered."""
    from oracle import oracle
    from revenant_amd import _lib
    from revenant_amd.delivery import Delivery
    from revenant_amd.dsp import build_spectrum_window, build_twiddles

    codes = _stream()
    proto = design_prototype(GRID, beta=8.0)
    eng, ids = _engine(dev)
    n = 256
    dl = Delivery(eng, RATE, transform=n, group_delay=proto.group_delay_samples)
    assert dl.spectrum_frame(4096) is None
    frames = []
    for at in range(0, SAMPLES, 65536):
        eng.push(codes[2 * at:2 * (at + 65536)].view(np.uint32))
        print(f"delivery frame after {at + 65536} samples", flush=True)
        f = dl.spectrum_frame(4096)
        if f is not None:
            frames.append(f)
    assert len(frames) >= 6 and [f.sequence for f in frames] == list(range(len(frames)))
    f = frames[-1]
    st = eng.stats()
    first = int(st.next_block) - n
    assert f.start == first * GRID.decimation - proto.group_delay_samples
    assert f.count == n * GRID.decimation
    # the same frame from the oracle on the engine's ring, read back whole
    import ctypes as C
    m, rb = GRID.channels, 4096
    ring_h = np.zeros(2 * m * rb, np.float32)
    _lib.check(_lib.lib().rvx_buf_read(dev.ctx, ring_h.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(eng.channel_ring()), ring_h.nbytes))
    p = _lib.SpectrumParams(m, n, n.bit_length() - 1, rb, rb - 1, first & (rb - 1))
    want = oracle.spectrum(p, ring_h.view(np.complex64), build_twiddles(n),
                           build_spectrum_window(n, GRID.taps_per_branch))
    assert np.array_equal(f.power_db.view(np.uint32), want.view(np.uint32))
    lp = _lib.SpectrumLevelsParams(f.geometry.bins, 50, 990)
    lv = oracle.spectrum_levels(lp, want)
    assert (f.percentile_low_db, f.percentile_high_db) == (float(lv[0]), float(lv[1]))
    assert f.floor_db <= f.ceiling_db - 12.0 + 1e-3
    peak = int(np.argmax(f.power_db))
    nearest = int(round((200_000 - f.geometry.bin_zero_hz()) / f.geometry.bin_width_hz()))
    assert abs(peak - nearest) <= 1
    chunk = dl.audio_chunk(ids[0], 48_000)
    assert chunk.samples.size > 1000 and chunk.channels == 1
    assert chunk.source_end == int(st.next_block) * GRID.decimation - proto.group_delay_samples
    eng.close()
