"""GPU: the streaming engine (Graph::on_block + the receivers' stream cursors).

Blocking invariance, after tests/reference/test_vrx.cpp:1241 (a receiver's
output does not depend on how the stream is cut into blocks): the same stream
pushed in different cuttings gives bit-identical audio, averaged spectrum and
block counts. And the engine's outputs equal the oracle's composition over the
whole stream: the chain twin over every block, the averaged-spectrum twin over
every row, and each receiver's fine + demod twins over the audio the engine
produced. The SAM and DSB receivers take the fine ring and their carrier loop
between the two stages, the loop carrying its state from push to push; the
loop's output does not depend on where dispatches fall, so the twin runs it
in one dispatch from the engine's first fine sample.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, rng
from oracle import oracle
from revenant_amd.dsp import (CARRIER_NONE, DEMOD_AM, DEMOD_CW, DEMOD_DSB, DEMOD_NFM, DEMOD_SAM,
                              DEMOD_USB, DEMOD_WFM, FORMAT_CS16, GridParams,
                              VrxRequest, build_nco_table, build_twiddles, design_prototype,
                              ema_frame_decay, ema_params, plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
GRID = GridParams(64, 16, 32)   # 2x oversampled, 75 kHz channels
SAMPLES = 1 << 20
FRAME, ROWS = 4, 2
ALPHA = np.float32(1.0 / 16.0)
REQUESTS = [VrxRequest(200_000, DEMOD_NFM, 12500), VrxRequest(-333_333, DEMOD_AM, 10000),
            VrxRequest(510_000, DEMOD_USB, 2700), VrxRequest(-1_000_500, DEMOD_CW, 500,
                                                                 cw_pitch=700),
            VrxRequest(850_000, DEMOD_SAM, 10000), VrxRequest(-650_000, DEMOD_DSB, 6000),
            VrxRequest(-150_000, DEMOD_WFM, 60_000)]  # FM stereo, 75 us, as planned by default
STEREO = 6


@pytest.fixture(scope="module")
def stream():
    g = rng(0x454E47494E450001)
    codes = g.integers(-9000, 9000, size=2 * SAMPLES, dtype=np.int16)
    # a few strong tones so the receivers carry signal
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 6000), (-333_333, 4000), (510_400, 3000), (-1_000_500, 5000)):
        z = a * np.exp(2j * np.pi * f * t / RATE)
        codes[0::2] = np.clip(codes[0::2] + z.real, -32768, 32767).astype(np.int16)
        codes[1::2] = np.clip(codes[1::2] + z.imag, -32768, 32767).astype(np.int16)
    # AM on a carrier 120 Hz off the SAM receiver; DSB of a 700 Hz tone 50 Hz
    # off the DSB receiver
    tone = np.cos(2 * np.pi * 700.0 * t / RATE)
    for f, env in ((850_120, 5000 * (1.0 + 0.6 * tone)), (-649_950, 6000 * tone)):
        z = env * np.exp(2j * np.pi * f * t / RATE)
        codes[0::2] = np.clip(codes[0::2] + z.real, -32768, 32767).astype(np.int16)
        codes[1::2] = np.clip(codes[1::2] + z.imag, -32768, 32767).astype(np.int16)
    proto = design_prototype(GRID, beta=8.0).taps
    plans = [plan_vrx(GRID, RATE, r) for r in REQUESTS]
    # a pilot-tone stereo multiplex (1 kHz left, 3 kHz right) on the WFM
    # receiver, at its deviation
    wfm = plans[STEREO].plan
    left, right = np.sin(2 * np.pi * 1000.0 * t / RATE), np.sin(2 * np.pi * 3000.0 * t / RATE)
    wp = 2 * np.pi * 19_000.0 * t / RATE
    m = 0.4 * ((left + right) / 2 + (left - right) / 2 * np.sin(2 * wp)) + 0.1 * np.sin(wp)
    phase = 2 * np.pi * wfm.deviation * np.cumsum(m) / RATE
    z = 5000 * np.exp(1j * (2 * np.pi * -150_000 * t / RATE + phase))
    codes[0::2] = np.clip(codes[0::2] + z.real, -32768, 32767).astype(np.int16)
    codes[1::2] = np.clip(codes[1::2] + z.imag, -32768, 32767).astype(np.int16)
    return codes.view(np.uint32), proto, plans


def _run(dev, stream, cuts, on_device=False, start=0):
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096,
                 frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA), start_index=start)
    ids = [eng.add(p, audio_capacity=1 << 16) for p in plans]
    pos = 0
    src = torch.from_numpy(codes.view(np.int32)).cuda() if on_device else codes
    for n in cuts:
        eng.push(src[pos:pos + n])
        pos += n
    assert pos == codes.size
    audio = [eng.read_audio(i) for i in ids]
    spec = eng.spectrum()
    st = eng.stats()
    out = (audio, spec, (st.blocks, st.next_block, st.spectrum_rows, st.audio_lost))
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


def test_engine_blocking_invariance(dev, stream):
    even = _run(dev, stream, [65536] * (SAMPLES // 65536))
    ragged = _run(dev, stream, _cuts(7, SAMPLES, 1, 9000))
    tiny_then_big = _run(dev, stream, [1, 31, 32, 33, 1000] + _cuts(8, SAMPLES - 1097, 20000,
                                                                     65536), on_device=True)
    for other, name in ((ragged, "ragged"), (tiny_then_big, "tiny then big, device pushes")):
        assert other[2][:3] == even[2][:3], name
        assert_bit_exact(other[1], even[1], what=f"spectrum, {name}")
        for v, ((a0, x), (b0, y)) in enumerate(zip(even[0], other[0])):
            assert a0 == b0, f"receiver {v} first audio, {name}"
            assert x.size > 10000
            # a cut can leave the last few outputs for a later push; compare the common part
            n = min(x.size, y.size)
            assert n > x.size - 64
            assert_bit_exact(y[:n], x[:n], what=f"receiver {v} audio, {name}")


@pytest.fixture(scope="module")
def oracle_run(stream):
    """The chain twin over every block of the stream (from sample 0) and the
    averaged-spectrum twin over every row."""
    codes, proto, plans = stream
    M, D, N = GRID.channels, GRID.decimation, GRID.prototype_length
    first = (N - 1 + D - 1) // D
    blocks = (SAMPLES - 1) // D - first + 1
    blocks -= blocks % (FRAME * ROWS)
    cap = 1 << (blocks - 1).bit_length()
    od = oracle.OChain(FORMAT_CS16, M, GRID.taps_per_branch, D, first, blocks, cap, FRAME,
                       ALPHA, ema_frame_decay(ALPHA, FRAME))
    ring, partial = oracle.chain(od, codes.view(np.uint8), proto, build_twiddles(M), threads=4)
    state = np.zeros(M, dtype=np.float32)
    oracle.ema_combine(ema_params(M, cap, first, FRAME, blocks // FRAME, ALPHA, ROWS), partial,
                       state)
    return first, blocks, cap, ring.reshape(M, cap, 2), state


def _oracle_audio(p, ring, cap, a0, n):
    """Receiver p's audio a0 .. a0+n-1 from the twin ring; channel sample m
    sits in slot m & (cap-1), also for absolute indices past the stream's
    start when the start is a multiple of M (the output phase correction
    depends on m*D mod M only)."""
    r, hist = p.demod.decimation, p.history
    fine_count = (n - 1) * r + hist + 1
    out_cap = 1 << (fine_count - 1).bit_length()
    chan = np.ascontiguousarray(ring[p.plan.placement.channel]).ravel()
    fp, _ = p.fine_block(0, cap - 1, out_cap - 1, a0 * r - hist, fine_count)
    nco = build_nco_table(16)
    fine = oracle.vrx_fine(p.fine, fp, chan, p.fine_taps, nco, out_cap)
    if p.plan.carrier.power != CARRIER_NONE:
        state = np.zeros(8, dtype=np.uint32)
        cp = p.carrier_block(out_cap - 1, a0 * r - hist, fine_count, reset=True)
        oracle.vrx_carrier(p.plan.carrier, cp, nco, fine, state)
        assert state[5] == 1, "the loop locked"
    dp = p.demod_block(out_cap - 1, a0, n)
    return oracle.vrx_demod(p.demod, dp, fine, p.weights)


def test_engine_matches_oracle(dev, stream, oracle_run):
    codes, proto, plans = stream
    first, blocks_all, cap, ring, state = oracle_run
    (audio, spec, (blocks, next_block, rows, lost)) = _run(dev, stream, _cuts(9, SAMPLES, 5000,
                                                                              60000))
    assert lost == 0
    assert blocks == blocks_all and next_block == first + blocks
    assert rows == blocks // (FRAME * ROWS)
    assert_bit_exact(spec, state, what="averaged spectrum after the stream")
    for v, (p, (a0, got)) in enumerate(zip(plans, audio)):
        ref = _oracle_audio(p, ring, cap, a0, got.size // p.demod.channels)
        assert_bit_exact(got, ref, what=f"engine receiver {v} vs oracle")
        assert np.abs(got).max() > 0
    # The stereo receiver runs the stereo branch with its pilot gate open (L and
    # R differ). A 60 kHz receiver passes +-30 kHz and so not the 38 kHz
    # subcarrier: separation is checked where the whole station reaches the
    # demodulator (test_stereo.py, test_gpu_vrx.py).
    p, (a0, got) = plans[STEREO], audio[STEREO]
    assert p.demod.channels == 2 and p.demod.pilot_taps > 0
    lch, rch = got[0::2], got[1::2]
    assert np.count_nonzero(lch == rch) * 20 < lch.size


def test_engine_late_start_and_receiver_removal(dev, stream, oracle_run):
    """A stream starting past 2^32 (a multiple of M, so the channel values
    equal the twin's at shifted block indices) gives the twin's spectrum and
    audio; a receiver removed midway leaves the others unchanged."""
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    first, blocks_all, cap, ring, state = oracle_run
    start = 1 << 33
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096,
                 frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA), start_index=start)
    ids = [eng.add(p, audio_capacity=1 << 16) for p in plans]
    half = SAMPLES // 2
    for pos in range(0, half, 50000):
        eng.push(codes[pos:min(pos + 50000, half)])
    eng.remove(ids[1])
    for pos in range(half, SAMPLES, 50000):
        eng.push(codes[pos:min(pos + 50000, SAMPLES)])
    st = eng.stats()
    assert st.blocks == blocks_all and st.audio_lost == 0
    assert_bit_exact(eng.spectrum(), state, what="late-start spectrum")
    for i, p in [(ids[k], plans[k]) for k in (0, 2, 3, 4, 5)]:
        a0, got = eng.read_audio(i)
        assert a0 * p.demod.decimation * p.plan.channel_rate // p.plan.demod_rate > start // 32
        ref = _oracle_audio(p, ring, cap, a0, got.size // p.demod.channels)
        assert_bit_exact(got, ref, what=f"late start receiver {i}")
    eng.close()


def test_engine_refuses_small_ring(dev, stream):
    from revenant_amd import _lib
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    with pytest.raises(_lib.RvxError):
        Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=1024)


def test_engine_front_end_correction_removes_spike_and_image(dev):
    """After test_front_end_correction.cpp:262-324 (the engine's stage on the
    device): a cs16 stream with a -33 dBFS DC offset, a 1.05 gain ratio and a
    3 degree phase error around a 0.1 tone 300 kHz up. With calibration on,
    the averaged spectrum's DC channel drops by 30 dB and the image channel
    (300 kHz down) by 25 dB against the tone, and the applied correction is
    the model's inverse."""
    from revenant_amd.engine import CALIBRATE_DC, CALIBRATE_IQ, Engine

    n = 2 * RATE
    t = np.arange(n, dtype=np.float64)
    ang = 2 * np.pi * 300_000 * t / RATE
    g = rng(0x464545494D500001)
    phi = 3.0 * np.pi / 180.0
    i_r = 0.1 * np.cos(ang) + 0.02 + 1e-4 * g.standard_normal(n)
    q_r = 1.05 * (0.1 * np.sin(ang) * np.cos(phi) + 0.1 * np.cos(ang) * np.sin(phi)) + 0.01 + \
        1e-4 * g.standard_normal(n)
    codes = np.empty(2 * n, dtype=np.int16)
    codes[0::2] = np.clip(np.round(i_r * 32768), -32768, 32767)
    codes[1::2] = np.clip(np.round(q_r * 32768), -32768, 32767)
    codes = codes.view(np.uint32)
    proto = design_prototype(GRID, beta=8.0).taps

    def run(cal):
        eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=8192,
                     frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA), calibration=cal,
                     sample_rate=RATE)
        for pos in range(0, n, 65536):
            eng.push(codes[pos:pos + 65536])
        spec = eng.spectrum().astype(np.float64)
        est, applied = eng.front_end()
        eng.close()
        return spec, est, applied

    off, est_off, applied_off = run(0)
    on, est, applied = run(CALIBRATE_DC | CALIBRATE_IQ)
    M = GRID.channels
    tone, image = 300_000 * M // RATE, (-300_000 * M // RATE) % M
    db = lambda x: 10 * np.log10(max(x, 1e-30))
    assert not est_off.measured and (applied_off.cross, applied_off.scale) == (0.0, 1.0)
    assert est.measured and est.iq_plausible
    dc_drop = db(off[0]) - db(on[0])
    irr_off = db(off[tone]) - db(off[image])
    irr_on = db(on[tone]) - db(on[image])
    print(f"front end: DC drop {dc_drop:.1f} dB, image rejection {irr_off:.1f} -> {irr_on:.1f} dB",
          flush=True)
    assert abs(irr_off - 28.9) < 0.5, irr_off
    assert dc_drop >= 30.0, dc_drop
    assert irr_on - irr_off >= 25.0, (irr_off, irr_on)
    assert abs(applied.cross - (-np.tan(phi))) < 2e-3
    assert abs(applied.scale - 1.0 / (1.05 * np.cos(phi))) < 2e-3
    assert abs(applied.dc_i - 0.02) < 1e-3 and abs(applied.dc_q - 0.01) < 1e-3


def test_engine_blanker_feeds_the_fine_stage(dev, stream, oracle_run):
    """Impulses on two receivers' channels: with the blanker on, each
    receiver's audio equals the blanker twins over its whole channel (zeros
    below the first block, as a cleared ring reads) followed by the fine and
    demod twins over the blanked channel; and it differs from the unblanked
    audio where the impulses were."""
    from revenant_amd import _lib
    from revenant_amd.dsp import BLANK_APPLY, BLANK_DETECT, blanker_threshold, design_blanker
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    first, blocks_all, cap, ring, state = oracle_run
    # impulses: a few strong clicks in the raw stream (every channel sees them)
    g = rng(0x424C414E4B)
    clicked = codes.copy().view(np.int16)
    for at in g.integers(100_000, SAMPLES - 100_000, size=12):
        clicked[2 * at:2 * at + 32] = 30000  # 16 samples
    clicked = clicked.view(np.uint32)
    od = oracle.OChain(FORMAT_CS16, GRID.channels, GRID.taps_per_branch, GRID.decimation, first,
                       blocks_all, cap, FRAME, ALPHA, ema_frame_decay(ALPHA, FRAME))
    cring, _ = oracle.chain(od, clicked.view(np.uint8), proto, build_twiddles(GRID.channels),
                            threads=4)
    cring = cring.reshape(GRID.channels, cap, 2)
    sel = [0, 1]  # an NFM and an AM receiver
    cfg = design_blanker(plans[0].plan.channel_rate)
    # 6 dB: a click the channelizer has smeared over its impulse response
    # seldom stands 12 dB over its own window
    thr = blanker_threshold(6.0)
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096,
                 frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA))
    ids = [eng.add(plans[k], audio_capacity=1 << 16) for k in sel]
    for rid in ids:
        eng.set_blanker(rid, cfg, thr)
    for pos in range(0, SAMPLES, 60000):
        eng.push(clicked[pos:min(pos + 60000, SAMPLES)])
    assert eng.stats().audio_lost == 0
    for k, rid in zip(sel, ids):
        p = plans[k]
        a0, got = eng.read_audio(rid)
        # the twin: the channel linear from `first`, zeros below it
        lin = 1 << 16
        P = 1024
        chan = np.zeros((lin, 2), np.float32)
        chan[P:P + blocks_all] = cring[p.plan.placement.channel][
            (first + np.arange(blocks_all)) & (cap - 1)]
        flags = np.zeros(lin, np.float32)
        blanked = np.zeros((lin, 2), np.float32)
        bp = _lib.BlankerParams(0, lin - 1, P, lin - 1, P, blocks_all, thr)
        c = _lib.BlankerConfig.from_buffer_copy(bytes(cfg))
        c.pass_ = BLANK_DETECT
        oracle.blank_detect(c, bp, chan, flags)
        ap = _lib.BlankerParams(0, lin - 1, P, lin - 1, P, blocks_all - cfg.lead, thr)
        c.pass_ = BLANK_APPLY
        oracle.blank_apply(c, ap, chan, flags, blanked)
        assert (flags != 0).sum() >= 1
        assert not np.array_equal(blanked[P:P + blocks_all - cfg.lead],
                                  chan[P:P + blocks_all - cfg.lead])
        # back onto the twin ring's slots, so the fine/demod composition reads it
        bring = np.zeros((GRID.channels, cap, 2), np.float32)
        m = first + np.arange(blocks_all - cfg.lead)
        bring[p.plan.placement.channel][m & (cap - 1)] = blanked[P:P + blocks_all - cfg.lead]
        ref = _oracle_audio(p, bring, cap, a0, got.size // p.demod.channels)
        assert_bit_exact(got, ref, what=f"blanked receiver {k}")
    eng.close()


def test_engine_display_tap_and_pane(dev, stream, oracle_run):
    """A receiver's display tap in the engine: after the stream, its pane is
    the spectrum twin (two channels, unity gains, channel 0's half) over the
    newest 1024 display samples, and those samples are the fine twin with
    the display table over the twin's channel, bit for bit."""
    from revenant_amd import _lib
    from revenant_amd.dsp import build_spectrum_window, plan_vrx_display
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    first, blocks_all, cap, ring, state = oracle_run
    p = plans[0]  # NFM 12.5 kHz
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096,
                 frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA))
    rid = eng.add(p, audio_capacity=1 << 16)
    n, dcap = 1024, 1 << 14
    eng.set_display(rid, dcap, n)
    assert eng.passband(rid, n) is None
    for pos in range(0, SAMPLES, 50000):
        eng.push(codes[pos:min(pos + 50000, SAMPLES)])
    pane, end, rate = eng.passband(rid, n)
    d = plan_vrx_display(p)
    assert rate == d.plan.display_rate and end * d.plan.decimation <= first + blocks_all
    chan = np.ascontiguousarray(ring[p.plan.placement.channel]).ravel()
    fp, _ = d.block(0, cap - 1, dcap - 1, end - n, n)
    disp = oracle.vrx_fine(d.plan.fine, fp, chan, d.taps, build_nco_table(16), dcap)
    two = np.zeros(4 * dcap, np.float32)
    two[:2 * dcap] = disp
    win = build_spectrum_window(n, GRID.taps_per_branch).copy()
    win[n:n + n // 2] = 1.0
    sp = _lib.SpectrumParams(2, n, 10, dcap, dcap - 1, (end - n) & (dcap - 1))
    want = oracle.spectrum(sp, two, build_twiddles(n), win)[n // 2:]
    assert_bit_exact(pane, want, what="engine passband pane")
    # the receiver's own tone (200 kHz sits at the receiver's centre) is the pane's peak
    assert abs(int(np.argmax(pane)) - n // 4) <= 2
    eng.close()


def test_a_lagging_reader_loses_the_oldest_audio_and_it_is_counted(dev, stream):
    """Audio the reader did not copy before the ring wrapped over it is
    counted in audio_overrun; the reader resumes at the oldest output held,
    which is bit-identical to the same output read in time."""
    from revenant_amd.engine import Engine

    codes, proto, plans = stream
    plan = plans[0]  # NFM
    cap = 1 << 12
    runs = {}
    for lag in (False, True):
        eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=65536, ring_blocks=4096,
                     frame_blocks=FRAME, row_frames=ROWS, alpha=float(ALPHA))
        rid = eng.add(plan, audio_capacity=cap)
        got, pos = [], 0
        for n in [65536] * (SAMPLES // 65536):
            eng.push(codes[pos:pos + n])
            pos += n
            if not lag:
                got.append(eng.read_audio(rid))
        if lag:
            got.append(eng.read_audio(rid))
        st = eng.stats()
        runs[lag] = (got, st.audio_overrun, st.audio_lost)
        eng.close()
    in_time, overrun, lost = runs[False]
    assert overrun == 0 and lost == 0
    produced = sum(x.size for _, x in in_time)
    (first, tail), = runs[True][0]
    assert tail.size == cap
    assert runs[True][1] == produced - cap
    assert first == in_time[0][0] + produced - cap
    whole = np.concatenate([x for _, x in in_time])
    assert_bit_exact(tail, whole[-cap:], what="the audio the lagging reader still gets")
