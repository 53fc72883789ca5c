"""GPU: the comb canceller end to end through the engine (the reference's
tests/engine/test_engine_comb.cpp:199): a drifting comb of 41 lines over white
noise pushed into the engine as cs16 codes, the comb detected on the last
second's spectrum of the input once a second (the reference detects on its
own full-span spectrum; here the test computes it from the same samples),
the bank planning every push's lines for rvx_engine_set_comb and reading back
rvx_engine_comb_projections. The stage must lock and take the comb out of
what the channelizer sees: each channel's power over the last second against
the same engine fed the noise alone."""
import ctypes as C

import numpy as np
import pytest

import comb_loop as cl
from revenant_amd import _lib
from revenant_amd.comb import CombCanceller
from revenant_amd.dsp import FORMAT_CS16, GridParams, design_prototype

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SECONDS = 4.0
SCALE = 16384.0
GRID = GridParams(16, 17, 16)
PUSH = 65536
RING = 1 << 17


def codes_of(x: np.ndarray) -> np.ndarray:
    iq = np.empty(2 * x.size, dtype=np.float64)
    iq[0::2], iq[1::2] = x.real * SCALE, x.imag * SCALE
    return np.clip(np.round(iq), -32768, 32767).astype(np.int16).view(np.uint32)


def channel_power(dev, eng, last_blocks: int) -> np.ndarray:
    """Mean |x|^2 of each channel over the engine's last `last_blocks` blocks."""
    pitch = eng.ring_pitch
    host = np.zeros(2 * GRID.channels * pitch, dtype=np.float32)
    _lib.check(dev.lib.rvx_buf_read(dev.ctx, host.ctypes.data, C.c_void_p(eng.channel_ring()),
                                    host.nbytes))
    ring = host.view(np.complex64).reshape(GRID.channels, pitch)[:, :RING]
    nxt = eng.stats().next_block
    slots = np.arange(nxt - last_blocks, nxt) & (RING - 1)
    return np.mean(np.abs(ring[:, slots].astype(np.complex128)) ** 2, axis=1)


def run_engine(dev, codes: np.ndarray, comb_on: bool, scene: np.ndarray | None = None):
    from revenant_amd.engine import Engine

    eng = Engine(dev, GRID, design_prototype(GRID, beta=8.0).taps, FORMAT_CS16, max_push=PUSH,
                 ring_blocks=RING)
    canc = CombCanceller(eng, cl.RATE) if comb_on else None
    found = []
    next_detect = cl.RATE
    for start in range(0, codes.size - PUSH + 1, PUSH):
        if canc is not None:
            if start >= next_detect:
                found.append(len(canc.detect(cl.spectrum(scene[start - cl.RATE:start], 16384))))
                next_detect += cl.RATE
            canc.push(codes[start:start + PUSH], PUSH)
        else:
            eng.push(codes[start:start + PUSH])
    power = channel_power(dev, eng, cl.RATE // GRID.decimation)
    status = canc.status() if canc is not None else None
    eng.close()
    return power, status, found


def test_engine_comb_stage_locks_and_takes_the_comb_out(dev):
    scene = cl.render(SECONDS)
    noise = cl.render(SECONDS, lines=False)   # the same noise, drawn first from the seed
    codes = codes_of(scene)
    print("comb through the engine: pushing", codes.size, "samples three times", flush=True)
    p_off, _, _ = run_engine(dev, codes, False)
    p_noise, _, _ = run_engine(dev, codes_of(noise), False)
    p_on, st, found = run_engine(dev, codes, True, scene=codes.view(np.int16).astype(
        np.float32).view(np.complex64) / SCALE)
    off_db = 10 * np.log10(p_off / p_noise)
    on_db = 10 * np.log10(p_on / p_noise)
    print(f"detections {found}; lock {st.lock} lines {st.lines_cancelled} suppression "
          f"{st.suppression_db:.1f} dB spacing {st.comb.spacing_hz:.3f} Hz; channel power over "
          f"the noise alone: off {np.round(off_db, 2)} on {np.round(on_db, 2)}")
    assert found and found[0] > 0
    assert st.lock == _lib.COMB_LOCKED and st.lines_cancelled >= _lib.COMB_MIN_LINES
    held = off_db > 3.0                         # channels the comb stands out in
    assert held.sum() >= 8
    assert np.all(on_db[held] < 0.5)            # back to the noise's own power
    assert np.mean(off_db[held] - on_db[held]) > 4.0
