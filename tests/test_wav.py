"""IQ recordings in WAV containers (revenant_amd/wav.py): the reference's
container cases (tests/engine/test_sources.cpp:1084-1330) and the rest of the
format's rules, on files built here byte by byte."""
import struct

import numpy as np
import pytest

from revenant_amd.capture import CaptureError
from revenant_amd.wav import WavFile, is_wav, read_wav_metadata, write_wav


def _ramp_cf32(n, seed):
    """Distinct float pairs, so an offset by any header byte shows."""
    k = np.arange(n, dtype=np.float32)
    out = np.empty(2 * n, dtype=np.float32)
    out[0::2] = k * np.float32(1e-3) + np.float32(seed)
    out[1::2] = -k * np.float32(2e-3) - np.float32(seed)
    return out


def _riff(chunks, signature=b"RIFF", form=b"WAVE"):
    """A container from (id, payload, declared size or None) chunks, padded."""
    body = b""
    for cid, data, declared in chunks:
        size = len(data) if declared is None else declared
        body += cid + struct.pack("<I", size) + data + (b"\0" if len(data) & 1 else b"")
    size = 0xFFFFFFFF if signature in (b"RF64", b"BW64") else 4 + len(body)
    pack = ">I" if signature == b"RIFX" else "<I"
    return signature + struct.pack(pack, size) + form + body


def _fmt(tag=3, channels=2, rate=2_000_000, bits=32, align=None):
    align = channels * bits // 8 if align is None else align
    return (b"fmt ", struct.pack("<HHIIHH", tag, channels, rate, rate * align, align, bits), None)


def _auxi(centre, rate, when=None):
    st = struct.pack("<8H", *when) if when else bytes(16)
    return (b"auxi", st + bytes(16) + struct.pack("<II", centre, rate) + bytes(20), None)


def test_auxi_supplies_rate_centre_and_start(tmp_path):
    path = str(tmp_path / "HDSDR_20260921_140307Z_7100kHz_RF.wav")
    x = _ramp_cf32(12_000, 11)
    open(path, "wb").write(_riff([_fmt(rate=1_000_000),
                                  _auxi(7_100_000, 1_000_000, (2026, 9, 1, 21, 14, 3, 7, 250)),
                                  (b"data", x.tobytes(), None)]))
    m = read_wav_metadata(path)
    assert (m.rate, m.center_hz, m.format, m.total_samples) == (1_000_000, 7_100_000, "cf32",
                                                                12_000)
    import calendar
    assert m.anchor_ns == calendar.timegm((2026, 9, 21, 14, 3, 7)) * 10**9 + 250 * 10**6
    f = WavFile(path)
    got = np.concatenate([b.codes for b in f.blocks(2048)])
    assert got[0] == x[0] and got[-1] == x[-1] and np.array_equal(got, x)


def test_a_data_chunk_claiming_more_than_the_file_is_refused(tmp_path):
    path = str(tmp_path / "cut.wav")
    payload = _ramp_cf32(1000, 0).tobytes()
    open(path, "wb").write(_riff([_fmt(), (b"data", payload, 4 * len(payload))]))
    with pytest.raises(CaptureError, match=rf"data chunk.*{4 * len(payload)}"):
        read_wav_metadata(path)


@pytest.mark.parametrize("signature", ["RF64", "BW64"])
def test_wide_containers_take_the_data_size_from_ds64(tmp_path, signature):
    path = str(tmp_path / "long.wav")
    x = _ramp_cf32(6000, 5)
    write_wav(path, x, "cf32", 2_000_000, container=signature)
    m = read_wav_metadata(path)
    assert m.container == signature and m.total_samples == 6000
    got = WavFile(path, center_hz=10_000_000).codes(0, 6000)
    assert np.array_equal(got, x)


def test_a_ds64_size_the_file_cannot_hold_is_refused(tmp_path):
    path = str(tmp_path / "liar.wav")
    payload = _ramp_cf32(500, 0).tobytes()
    ds64 = struct.pack("<QQQI", 0, 5_000_000_000, 0, 0)
    open(path, "wb").write(_riff([(b"ds64", ds64, None), _fmt(),
                                  (b"data", payload, 0xFFFFFFFF)], signature=b"BW64"))
    with pytest.raises(CaptureError, match="5000000000"):
        read_wav_metadata(path)


def test_the_sentinel_without_ds64_is_refused(tmp_path):
    path = str(tmp_path / "sentinel.wav")
    open(path, "wb").write(_riff([_fmt(), (b"data", _ramp_cf32(500, 0).tobytes(), 0xFFFFFFFF)]))
    with pytest.raises(CaptureError, match="4294967295"):
        read_wav_metadata(path)


def test_rifx_is_refused_by_name(tmp_path):
    path = str(tmp_path / "swapped.wav")
    open(path, "wb").write(_riff([_fmt(), (b"data", _ramp_cf32(500, 0).tobytes(), None)],
                                 signature=b"RIFX"))
    assert is_wav(path)
    with pytest.raises(CaptureError, match="RIFX.*byte-swapped"):
        read_wav_metadata(path)


def test_no_centre_is_said_not_presented_as_dc(tmp_path):
    path = str(tmp_path / "perseus_like.wav")
    open(path, "wb").write(_riff([_fmt(), (b"data", _ramp_cf32(2000, 0).tobytes(), None)]))
    m = read_wav_metadata(path)
    assert m.center_hz is None and m.rate == 2_000_000
    assert any("no auxi chunk" in n and "center_hz" in n for n in m.notes)
    zero = str(tmp_path / "zero_centre.wav")
    open(zero, "wb").write(_riff([_fmt(), _auxi(0, 2_000_000),
                                  (b"data", _ramp_cf32(10, 0).tobytes(), None)]))
    m = read_wav_metadata(zero)
    assert m.center_hz is None and any("zero" in n for n in m.notes)


def test_an_auxi_rate_contradicting_fmt_is_refused(tmp_path):
    path = str(tmp_path / "clash.wav")
    open(path, "wb").write(_riff([_fmt(), _auxi(7_100_000, 192_000),
                                  (b"data", _ramp_cf32(1000, 0).tobytes(), None)]))
    with pytest.raises(CaptureError, match=r"2000000.*192000"):
        read_wav_metadata(path)


def test_a_channel_count_other_than_two_is_refused(tmp_path):
    path = str(tmp_path / "mono.wav")
    open(path, "wb").write(_riff([_fmt(channels=1, rate=48_000),
                                  (b"data", _ramp_cf32(1000, 0).tobytes(), None)]))
    with pytest.raises(CaptureError, match="1 channels"):
        read_wav_metadata(path)


@pytest.mark.parametrize("fmt,dtype", [("cu8", np.uint8), ("cs16", np.int16), ("cs24", np.uint8),
                                       ("cf32", np.float32)])
@pytest.mark.parametrize("extensible", [False, True])
def test_sample_formats_round_trip(tmp_path, fmt, dtype, extensible):
    g = np.random.default_rng(3)
    n = 777
    bps = {"cu8": 2, "cs16": 4, "cs24": 6, "cf32": 8}[fmt]
    raw = g.integers(0, 256, size=n * bps, dtype=np.uint8)
    if fmt == "cf32":
        raw = g.standard_normal(2 * n).astype(np.float32).view(np.uint8)
    path = str(tmp_path / f"{fmt}.wav")
    write_wav(path, raw, fmt, 250_000, center_hz=14_074_000, start_ns=1_700_000_000_123_000_000,
              extensible=extensible)
    m = read_wav_metadata(path)
    assert (m.format, m.rate, m.center_hz, m.total_samples) == (fmt, 250_000, 14_074_000, n)
    assert m.anchor_ns == 1_700_000_000_123_000_000
    got = WavFile(path).codes(0, n)
    assert got.dtype == dtype
    assert np.array_equal(got.view(np.uint8), raw.view(np.uint8))


def test_odd_chunks_are_padded_and_unknown_ones_skipped(tmp_path):
    """The pad byte after an odd payload is not counted in its size; a walk
    that forgets it reads every later chunk id one byte late."""
    path = str(tmp_path / "odd.wav")
    x = _ramp_cf32(64, 2)
    open(path, "wb").write(_riff([(b"LIST", b"INFOISFT\x05\x00\x00\x00abcde", None), _fmt(),
                                  (b"junk", b"\x01\x02\x03", None),
                                  (b"data", x.tobytes(), None)]))
    assert np.array_equal(WavFile(path, center_hz=1).codes(0, 64), x)


@pytest.mark.parametrize("case,match", [
    ("tag", "format tag 1 at 12 bits"), ("align", "block align"), ("nofmt", "no fmt"),
    ("nodata", "no data"), ("form", "not WAVE"), ("sig", "not RIFF, RF64 or BW64"),
    ("short", "12-byte")])
def test_refusals(tmp_path, case, match):
    path = str(tmp_path / f"{case}.wav")
    data = (b"data", _ramp_cf32(8, 0).tobytes(), None)
    blob = {
        "tag": _riff([_fmt(tag=1, bits=12), data]),
        "align": _riff([_fmt(align=6), data]),
        "nofmt": _riff([data]),
        "nodata": _riff([_fmt()]),
        "form": _riff([_fmt(), data], form=b"AVI "),
        "sig": _riff([_fmt(), data], signature=b"FORM"),
        "short": b"RIFF\x00\x00",
    }[case]
    open(path, "wb").write(blob)
    with pytest.raises(CaptureError, match=match):
        read_wav_metadata(path)


def test_riff_past_4_gib_is_refused(tmp_path):
    """A sparse file one byte past the 32-bit sizes under a RIFF signature."""
    path = tmp_path / "huge.wav"
    path.write_bytes(_riff([_fmt(), (b"data", _ramp_cf32(8, 0).tobytes(), None)]))
    with open(path, "r+b") as f:
        f.truncate((1 << 32) + 1)
    with pytest.raises(CaptureError, match="RF64 or BW64"):
        read_wav_metadata(str(path))


def test_rate_and_centre_arguments_must_agree_with_the_file(tmp_path):
    path = str(tmp_path / "agree.wav")
    write_wav(path, _ramp_cf32(16, 0), "cf32", 96_000, center_hz=7_000_000)
    with pytest.raises(CaptureError, match="rate=48000"):
        WavFile(path, rate=48_000)
    with pytest.raises(CaptureError, match="center=7100000"):
        WavFile(path, center_hz=7_100_000)
    assert WavFile(path, rate=96_000, center_hz=7_000_000).length == 16
