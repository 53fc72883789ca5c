"""GPU: the engine runs a receiver's audio noise stages on every push (the
reference records the line kernel and spectral reduction after each
receiver's demodulator). With the audio ring small enough that pushes wrap
it, every receiver's audio equals the oracle's line twin then spectral twin
run once over the same receiver's audio from an engine without the stages,
bit for bit: the stages' state carries across pushes and ring wraps exactly
as one dispatch over the whole stream would."""
import numpy as np
import pytest

from helpers import assert_bit_exact, rng
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, DEMOD_USB, FORMAT_CS16, GridParams, VrxRequest,
                              design_prototype, line_state_size, plan_vrx, spectral_frame_for,
                              spectral_state_size, spectral_tables)
from revenant_amd.noise import NoiseRequest, plan_noise

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000
GRID = GridParams(64, 16, 32)
SAMPLES = 1 << 19
PUSH = 65536
CAP = 3000


def _stream():
    g = rng(0x4E4F49534545)
    codes = g.integers(-6000, 6000, size=2 * SAMPLES, dtype=np.int16)
    t = np.arange(SAMPLES)
    for f, a in ((200_000, 6000), (-333_333, 4000), (510_400, 3000), (511_300, 2500)):
        z = a * np.exp(2j * np.pi * f * t / RATE)
        codes[0::2] = np.clip(codes[0::2] + z.real, -32768, 32767).astype(np.int16)
        codes[1::2] = np.clip(codes[1::2] + z.imag, -32768, 32767).astype(np.int16)
    return codes


def _receivers():
    usb = plan_vrx(GRID, RATE, VrxRequest(510_000, DEMOD_USB, passband=(100, 2_700)))
    nfm = plan_vrx(GRID, RATE, VrxRequest(200_000, DEMOD_NFM, 12_500))
    am = plan_vrx(GRID, RATE, VrxRequest(-333_333, DEMOD_AM, 10_000))
    reqs = [NoiseRequest(demod=DEMOD_USB, notch_enabled=True, notch_hz=1_300,
                         auto_notch_enabled=True, nr_enabled=True, nr_strength=0.5),
            NoiseRequest(demod=DEMOD_NFM, auto_notch_enabled=True),
            NoiseRequest(demod=DEMOD_AM, nr_enabled=True, nr_strength=0.7)]
    return [(p, plan_noise(r, p)) for p, r in zip((usb, nfm, am), reqs)]


def _run(dev, codes, receivers, noise):
    from revenant_amd.engine import Engine
    proto = design_prototype(GRID, beta=8.0).taps
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=PUSH, ring_blocks=4096)
    ids = [eng.add(p, audio_capacity=CAP) for p, _ in receivers]
    frame = spectral_frame_for(48_000)
    if noise:
        for rid, (_, np_) in zip(ids, receivers):
            eng.set_noise_plan(rid, np_, spectral_frame=frame)
    out = [[] for _ in ids]
    firsts = [None] * len(ids)
    for at in range(0, SAMPLES, PUSH):
        eng.push(codes[2 * at:2 * (at + PUSH)].view(np.uint32))
        for k, rid in enumerate(ids):
            first, a = eng.read_audio(rid)
            if a.size:
                if firsts[k] is None:
                    firsts[k] = first
                out[k].append(a)
    assert eng.stats().audio_lost == 0
    eng.close()
    return firsts, [np.concatenate(o) for o in out]


def test_engine_noise_stages_equal_the_twins_over_the_whole_stream(dev):
    codes = _stream()
    receivers = _receivers()
    print("engine without the noise stages", flush=True)
    f_plain, plain = _run(dev, codes, receivers, noise=False)
    print("engine with the noise stages", flush=True)
    f_noisy, noisy = _run(dev, codes, receivers, noise=True)
    frame = spectral_frame_for(48_000)
    tables = spectral_tables(frame)
    for k, (_, np_) in enumerate(receivers):
        assert f_plain[k] == f_noisy[k]
        assert plain[k].size == noisy[k].size and plain[k].size > 3 * CAP  # the ring wrapped
        want = plain[k].copy()
        if np_.line:
            c = _lib.LineConfig()
            lp = _lib.LineParams.from_buffer_copy(bytes(np_.line_params))
            lp.count = want.size
            oracle.noise_line(c, lp, want, np.zeros(line_state_size(c), np.float32))
        if np_.spectral:
            sp = _lib.SpectralParams.from_buffer_copy(bytes(np_.spectral_params))
            sp.count = want.size
            oracle.noise_spectral(frame, sp, tables, want,
                                  np.zeros(spectral_state_size(frame), np.float32))
        assert not np.array_equal(want, plain[k])
        assert_bit_exact(noisy[k], want, k, f"engine noise receiver {k}")


def test_set_noise_refuses_what_it_cannot_run(dev):
    from revenant_amd.engine import Engine
    proto = design_prototype(GRID, beta=8.0).taps
    eng = Engine(dev, GRID, proto, FORMAT_CS16, max_push=PUSH, ring_blocks=4096)
    rid = eng.add(plan_vrx(GRID, RATE, VrxRequest(200_000, DEMOD_NFM, 12_500)), audio_capacity=CAP)
    bad = _lib.LineParams(flags=2, stride=0)
    with pytest.raises(_lib.RvxError):
        eng.set_noise(rid, _lib.LineConfig(), bad)
    with pytest.raises(_lib.RvxError):
        eng.set_noise(rid + 1, _lib.LineConfig(), _lib.LineParams(flags=1))
    sp = _lib.SpectralParams()
    sp.eps = 1e-12
    with pytest.raises(_lib.RvxError):
        eng.set_noise(rid, spectral_frame=96, spectral_params=sp)
    eng.set_noise(rid)  # both off
    eng.close()
