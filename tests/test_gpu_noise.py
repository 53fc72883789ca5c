"""GPU vs oracle, bit for bit: the impulse blanker's detect and apply passes
across a ring wrap, after tests/reference/test_noise.cpp:163-268, over counts
that leave partial and single-lane workgroups and over the shortest and
longest windows and a long guard."""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import BLANK_APPLY, BLANK_DETECT
from tests.blanker_cases import (OUT_RING, apply_params, config_144k, impulsive_channels,
                                 wrapping_params)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0x4E4F495345000001


def _with(c, **fields):
    out = _lib.BlankerConfig.from_buffer_copy(bytes(c))
    for k, v in fields.items():
        setattr(out, k, v)
    return out


@pytest.mark.parametrize("count,shape", [(1500, {}), (257, {}), (1, {}),
                                         (1500, {"window": 16}), (1500, {"window": 256}),
                                         (1200, {"guard": 300, "hang": 5, "lead": 9})])
def test_detect_and_apply_match_twins_through_a_wrap(dev, count, shape):
    c = _with(config_144k(), **shape)
    ch = impulsive_channels(SEED + count)
    p, first = wrapping_params(count)
    d_ch = to_dev(ch.reshape(-1))
    want = np.zeros(OUT_RING, np.float32)
    oracle.blank_detect(_with(c, pass_=BLANK_DETECT), p, ch, want)
    got = torch.zeros(OUT_RING, dtype=torch.float32, device="cuda")
    print(f"blank detect count {count} {shape}", flush=True)
    dev.blank_detect(c, p, d_ch, got)
    assert_bit_exact(from_dev(got), want, SEED, "blanker flags")
    if count == 1500 and not shape:
        assert want.sum() > 20
    if count <= c.hang + c.lead:
        return
    a = apply_params(p, first, c)
    w_blanked = np.full((OUT_RING, 2), 3.0, np.float32)
    oracle.blank_apply(_with(c, pass_=BLANK_APPLY), a, ch, want, w_blanked)
    g_blanked = torch.full((2 * OUT_RING,), 3.0, dtype=torch.float32, device="cuda")
    print(f"blank apply count {a.count}", flush=True)
    dev.blank_apply(c, a, d_ch, to_dev(want), g_blanked)
    assert_bit_exact(from_dev(g_blanked), w_blanked.reshape(-1), SEED, "blanked ring")


def test_refusals_launch_nothing(dev):
    c = config_144k()
    p, _ = wrapping_params()
    ch = to_dev(impulsive_channels(SEED).reshape(-1))
    short = torch.zeros(OUT_RING - 1, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.RvxError):
        dev.blank_detect(c, p, ch, short)
    bad = _lib.BlankerParams.from_buffer_copy(bytes(p))
    bad.chan_base = 3 * 4096  # base + ring runs past the channel buffer
    with pytest.raises(_lib.RvxError):
        dev.blank_detect(c, bad, ch, torch.zeros(OUT_RING, dtype=torch.float32, device="cuda"))


# ---------------------------------------------------------------------------
# The line kernel and spectral noise reduction, after
# tests/reference/test_noise.cpp:330-470 and :880-909: audio and state
# carried across dispatches of ragged lengths, every dispatch compared.
# ---------------------------------------------------------------------------

from revenant_amd.dsp import (LINE_ALE, LINE_GUARD, LINE_NOTCH, LINE_STATE_WEIGHTS,  # noqa: E402
                              line_check, line_notch_flags, line_state_size, spectral_frame_for,
                              spectral_params, spectral_state_size, spectral_tables)
from tests.noise_audio_cases import LENGTHS, line_params, voice_audio  # noqa: E402

LINE_CASES = {
    "notch": (LINE_NOTCH, 0x4E4F495345000003),
    "ale": (LINE_ALE, 0x4E4F495345000004),
    "notch+ale": (LINE_NOTCH | LINE_ALE, 0x4E4F495345000005),
    "cw-guard": (LINE_NOTCH | LINE_ALE | LINE_GUARD, 0x4E4F495345000009),
    "four-notches": (line_notch_flags(4), 0x4E4F49534500000A),
    "two-notches+ale": (line_notch_flags(2) | LINE_ALE, 0x4E4F49534500000B),
}


@pytest.mark.parametrize("case", list(LINE_CASES))
def test_line_matches_twin_across_dispatches(dev, case):
    flags, seed = LINE_CASES[case]
    c = _lib.LineConfig()
    p = line_params(flags)
    all_audio = voice_audio(20000, seed, 48000.0)
    o_state = np.zeros(line_state_size(c), np.float32)
    g_state = to_dev(o_state)
    at = 0
    for n in LENGTHS:
        chunk = np.concatenate([all_audio[at:at + n], np.zeros(1, np.float32)])
        p.count = n
        line_check(c, p)
        want = chunk.copy()
        oracle.noise_line(c, p, want, o_state)
        got = to_dev(chunk)
        print(f"noise line {case} dispatch {n} at {at}", flush=True)
        dev.noise_line(c, p, got, g_state)
        assert_bit_exact(from_dev(got), want, seed, f"line audio ({n} at {at})")
        assert_bit_exact(from_dev(g_state), o_state, seed, f"line state ({n} at {at})")
        at += n
    if flags & LINE_ALE:
        assert np.abs(o_state[LINE_STATE_WEIGHTS:LINE_STATE_WEIGHTS + 128]).sum() > 0


def _items(struct_type, records):
    blob = b"".join(bytes(r) for r in records)
    return torch.frombuffer(bytearray(blob), dtype=torch.uint8).cuda()


def test_line_batch_runs_each_receiver_as_alone(dev):
    c = _lib.LineConfig()
    cases = list(LINE_CASES.values())
    audio = [voice_audio(3000 + 97 * i, seed, 48000.0) for i, (_, seed) in enumerate(cases)]
    states = [np.random.default_rng(i).normal(0, 0.01, line_state_size(c)).astype(np.float32)
              for i in range(len(cases))]
    for st in states:
        st[4] = 1234.0  # a history position inside the ring
    wants, d_audio, d_state, records = [], [], [], []
    for i, (flags, seed) in enumerate(cases):
        p = line_params(flags)
        p.count = audio[i].size
        w_a, w_s = audio[i].copy(), states[i].copy()
        oracle.noise_line(c, p, w_a, w_s)
        wants.append((w_a, w_s))
        d_audio.append(to_dev(audio[i]))
        d_state.append(to_dev(states[i]))
        item = _lib.NoiseLineItem(p, d_audio[-1].data_ptr(), d_state[-1].data_ptr())
        _lib.check(_lib.lib().rvx_noise_line_item_check(C.byref(c), C.byref(item), audio[i].size,
                                                        states[i].size))
        records.append(item)
    items = _items(_lib.NoiseLineItem, records)
    print(f"noise line batch of {len(records)}", flush=True)
    dev.noise_line_batch(c, items, len(records))
    for i, (w_a, w_s) in enumerate(wants):
        assert_bit_exact(from_dev(d_audio[i]), w_a, i, f"batched line audio {i}")
        assert_bit_exact(from_dev(d_state[i]), w_s, i, f"batched line state {i}")


@pytest.mark.parametrize("rate", [48000, 8000])
def test_spectral_matches_twin_across_dispatches(dev, rate):
    n = spectral_frame_for(rate)
    seed = 0x4E4F495345000006 + rate
    tables = spectral_tables(n)
    d_tables = to_dev(tables)
    all_audio = voice_audio(20000, seed, float(rate))
    o_state = np.zeros(spectral_state_size(n), np.float32)
    g_state = to_dev(o_state)
    at = 0
    for k in LENGTHS:
        chunk = np.concatenate([all_audio[at:at + k], np.zeros(1, np.float32)])
        p = spectral_params(0.7)
        p.count = k
        want = chunk.copy()
        oracle.noise_spectral(n, p, tables, want, o_state)
        got = to_dev(chunk)
        print(f"noise spectral frame {n} dispatch {k} at {at}", flush=True)
        dev.noise_spectral(n, p, d_tables, got, g_state)
        assert_bit_exact(from_dev(got), want, seed, f"spectral audio ({k} at {at})")
        assert_bit_exact(from_dev(g_state), o_state, seed, f"spectral state ({k} at {at})")
        at += k
    assert np.abs(want).max() > 0


def test_spectral_batch_runs_each_receiver_as_alone(dev):
    n = 512
    tables = spectral_tables(n)
    wants, d_audio, d_state, records = [], [], [], []
    for i in range(5):
        a = voice_audio(1500 + 313 * i, 100 + i, 48000.0)
        st = np.zeros(spectral_state_size(n), np.float32)
        st[0] = float((37 * i) % (n // 2))  # samples into the hop
        p = spectral_params(0.25 * i)
        p.count = a.size
        w_a, w_s = a.copy(), st.copy()
        oracle.noise_spectral(n, p, tables, w_a, w_s)
        wants.append((w_a, w_s))
        d_audio.append(to_dev(a))
        d_state.append(to_dev(st))
        item = _lib.NoiseSpectralItem(p, (C.c_uint32 * 2)(), d_audio[-1].data_ptr(),
                                      d_state[-1].data_ptr())
        _lib.check(_lib.lib().rvx_noise_spectral_item_check(n, C.byref(item), a.size, st.size))
        records.append(item)
    items = _items(_lib.NoiseSpectralItem, records)
    print("noise spectral batch of 5", flush=True)
    dev.noise_spectral_batch(n, to_dev(tables), items, len(records))
    for i, (w_a, w_s) in enumerate(wants):
        assert_bit_exact(from_dev(d_audio[i]), w_a, i, f"batched spectral audio {i}")
        assert_bit_exact(from_dev(d_state[i]), w_s, i, f"batched spectral state {i}")


def test_line_matches_twin_in_a_shape_of_its_own(dev):
    """A config other than the default 32 lanes x 128 taps (which runs a
    specialised kernel): 16 lanes of 6 taps over a 1024-sample history."""
    c = _lib.LineConfig(lanes=16, taps=96, history=1024)
    p = line_params(LINE_NOTCH | LINE_ALE | LINE_GUARD)
    audio = voice_audio(5000, 77, 48000.0)
    o_state = np.zeros(line_state_size(c), np.float32)
    g_state = to_dev(o_state)
    at = 0
    for n in (1000, 1, 2999, 1000):
        p.count = n
        line_check(c, p)
        want = audio[at:at + n].copy()
        oracle.noise_line(c, p, want, o_state)
        got = to_dev(audio[at:at + n].copy())
        print(f"noise line 16x96 dispatch {n} at {at}", flush=True)
        dev.noise_line(c, p, got, g_state)
        assert_bit_exact(from_dev(got), want, 77, f"line audio ({n} at {at})")
        assert_bit_exact(from_dev(g_state), o_state, 77, f"line state ({n} at {at})")
        at += n
