"""IQ recordings in WAV containers: RIFF WAVE, and RF64 and BW64 for files
past 4 GiB (core/source/file_source.cpp:1415-1760, read_wav_metadata; the
container sniff :1794).

The layouts this reader follows:
* RIFF (Microsoft/IBM, Multimedia Programming Interface and Data
  Specifications 1.0): a 12-byte header (signature, 32-bit size, form type
  "WAVE"), then chunks of a four-character id, a 32-bit little-endian payload
  size and the payload, padded to an even length by a byte the size does not
  count.
* RF64 (EBU Tech 3306) and BW64 (ITU-R BS.2088): the same chunks under
  another signature, with 0xFFFFFFFF in the 32-bit sizes and the true 64-bit
  sizes in a "ds64" chunk first: RIFF size, data size and sample count
  (8 bytes each), then a table length (4).
* "fmt " is WAVEFORMATEX: format tag, channels, sample rate, byte rate,
  block align, bits per sample (16 bytes). WAVE_FORMAT_EXTENSIBLE (tag
  0xFFFE) carries the real tag in the first two bytes of its SubFormat GUID,
  24 bytes into the chunk.
* "auxi" is the SDR recorders' chunk, which no standard describes: two Win32
  SYSTEMTIME blocks (start, stop: year, month, weekday, day, hour, minute,
  second, millisecond as 16-bit fields), then the centre frequency in hertz
  at byte 32 and the sample rate at byte 36 (60 bytes in all). Its rate is
  checked against the fmt chunk's, its start is taken only as a real date,
  and a zero centre means none.

An IQ recording has two channels. 8-bit PCM is unsigned offset binary (cu8),
16- and 24-bit PCM signed (cs16, cs24), 32-bit IEEE float cf32.
"""
from __future__ import annotations

import calendar
import os
import struct
from dataclasses import dataclass, field
from typing import Iterator

import numpy as np

from .capture import (BYTES_PER_SAMPLE, RECORDER_CLOCK_ACCURACY_NS, CaptureError, SourceBlock,
                      _view_codes)

SIZE_SENTINEL = 0xFFFFFFFF
TAG_PCM, TAG_FLOAT, TAG_EXTENSIBLE = 1, 3, 0xFFFE
AUXI_CORE = 60
DS64_CORE = 28


@dataclass
class WavMetadata:
    container: str           # "RIFF", "RF64" or "BW64"
    data_path: str
    data_offset: int
    data_bytes: int
    format: str              # cu8, cs16, cs24, cf32
    rate: int | None
    center_hz: int | None = None
    anchor_ns: int | None = None
    anchor_accuracy_ns: int = 0
    notes: list[str] = field(default_factory=list)

    @property
    def total_samples(self) -> int:
        return self.data_bytes // BYTES_PER_SAMPLE[self.format]


def is_wav(path: str) -> bool:
    """The container sniff: the bytes, not the extension, decide."""
    try:
        with open(path, "rb") as f:
            head = f.read(12)
    except OSError:
        return False
    return len(head) == 12 and head[8:12] == b"WAVE" and \
        head[:4] in (b"RIFF", b"RF64", b"BW64", b"RIFX")


def _systemtime_ns(b: bytes, where: str) -> int | None:
    year, month, _weekday, day, hour, minute, second, ms = struct.unpack_from("<8H", b, 0)
    if year == 0:
        return None
    if not (1 <= month <= 12 and 1 <= day <= 31 and hour < 24 and minute < 60 and second < 61
            and ms < 1000):
        raise CaptureError(f"{where} is not a calendar time ({year}-{month}-{day} "
                           f"{hour}:{minute}:{second}.{ms})")
    try:
        secs = calendar.timegm((year, month, day, hour, minute, second, 0, 0, 0))
    except (ValueError, OverflowError) as exc:
        raise CaptureError(f"{where} is not a calendar time: {exc}") from None
    return secs * 1_000_000_000 + ms * 1_000_000


def read_wav_metadata(path: str) -> WavMetadata:
    """The chunk walk up to what the samples need: format, rate, where the data
    lies and how much of it the file really holds; refusals say why."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        def at(offset: int, n: int) -> bytes:
            f.seek(offset)
            return f.read(n)

        head = at(0, 12)
        if len(head) < 12:
            raise CaptureError(f"'{path}': {len(head)} bytes cannot hold a 12-byte RIFF header")
        sig, form = head[:4].decode("latin-1"), head[8:12].decode("latin-1")
        if sig == "RIFX":
            raise CaptureError(f"'{path}' is RIFX (big-endian RIFF): its fields and samples are "
                               "byte-swapped and are not read as little-endian; convert it")
        if sig not in ("RIFF", "RF64", "BW64"):
            raise CaptureError(f"'{path}' starts with {sig!r}: not RIFF, RF64 or BW64")
        if form != "WAVE":
            raise CaptureError(f"'{path}' is a {sig} file of form {form!r}, not WAVE")
        wide = sig != "RIFF"
        if not wide and size > SIZE_SENTINEL:
            raise CaptureError(f"'{path}' is {size} bytes under a RIFF signature, whose 32-bit "
                               "sizes stop at 4 GiB: its sizes have wrapped (long recordings "
                               "are RF64 or BW64)")
        notes: list[str] = []
        fmt = None
        data = None
        ds64_data = None
        auxi = None
        pos = 12
        while pos + 8 <= size:
            hdr = at(pos, 8)
            if len(hdr) < 8:
                raise CaptureError(f"'{path}' ends inside a chunk header at byte {pos}")
            cid = hdr[:4].decode("latin-1")
            declared = struct.unpack_from("<I", hdr, 4)[0]
            body = pos + 8
            length = declared
            if cid == "data" and declared == SIZE_SENTINEL and ds64_data is not None:
                length = ds64_data
            if cid == "ds64":
                if not wide:
                    notes.append(f"'{path}': a ds64 chunk under a RIFF signature; its 64-bit "
                                 "sizes are not needed")
                if length < DS64_CORE:
                    raise CaptureError(f"'{path}': ds64 of {length} bytes, short of the "
                                       f"{DS64_CORE} holding its three sizes and table length")
                p = at(body, DS64_CORE)
                if len(p) < DS64_CORE:
                    raise CaptureError(f"'{path}' ends inside its ds64 chunk")
                ds64_data = struct.unpack_from("<Q", p, 8)[0]
            elif cid == "fmt ":
                if length < 16:
                    raise CaptureError(f"'{path}': fmt of {length} bytes, short of WAVEFORMATEX's "
                                       "16")
                p = at(body, min(length, 40))
                if len(p) < 16:
                    raise CaptureError(f"'{path}' ends inside its fmt chunk")
                tag, channels, rate, _byte_rate, align, bits = struct.unpack_from("<HHIIHH", p, 0)
                if tag == TAG_EXTENSIBLE:
                    if len(p) < 26:
                        raise CaptureError(f"'{path}': an extensible fmt chunk of {len(p)} bytes "
                                           "ends before its SubFormat names the sample type")
                    tag = struct.unpack_from("<H", p, 24)[0]
                fmt = (tag, channels, rate, align, bits)
            elif cid == "auxi":
                if length < AUXI_CORE:
                    notes.append(f"'{path}': auxi chunk of {length} bytes (its fields need "
                                 f"{AUXI_CORE}); no centre or start time read from it")
                else:
                    p = at(body, AUXI_CORE)
                    if len(p) < AUXI_CORE:
                        raise CaptureError(f"'{path}' ends inside its auxi chunk")
                    auxi = (struct.unpack_from("<I", p, 32)[0], struct.unpack_from("<I", p, 36)[0],
                            _systemtime_ns(p[:16], f"the auxi start time of '{path}'"))
            elif cid == "data":
                data = (body, length)
            step = length + (length & 1)  # the pad byte after an odd payload
            if step > size - body:
                break
            pos = body + step
    if fmt is None:
        raise CaptureError(f"'{path}' has no fmt chunk: nothing states the sample type")
    if data is None:
        raise CaptureError(f"'{path}' has no data chunk")
    offset, nbytes = data
    if nbytes > size - offset:
        raise CaptureError(f"'{path}': the data chunk at byte {offset} declares {nbytes} bytes "
                           f"but {size - offset} follow it; the file is cut short or the header "
                           "is not this file's, and what is there is not read as samples")
    tag, channels, rate, align, bits = fmt
    if channels != 2:
        raise CaptureError(f"'{path}': {channels} channels; an IQ recording has two (I, Q)")
    kinds = {(TAG_PCM, 8): "cu8", (TAG_PCM, 16): "cs16", (TAG_PCM, 24): "cs24",
             (TAG_FLOAT, 32): "cf32"}
    if (tag, bits) not in kinds:
        raise CaptureError(f"'{path}': format tag {tag} at {bits} bits is not read (8, 16 and "
                           "24-bit PCM and 32-bit float are)")
    want_align = channels * (bits // 8)
    if align not in (0, want_align):
        raise CaptureError(f"'{path}': block align {align} where {channels} channels of {bits} "
                           f"bits make {want_align}; the frame stride is in doubt")
    out = WavMetadata(sig, path, offset, nbytes, kinds[(tag, bits)], rate if rate > 0 else None,
                      notes=notes)
    if auxi is None:
        out.notes.append(f"'{path}' has no auxi chunk: no centre frequency in it (give "
                         "center_hz=)")
        return out
    centre, auxi_rate, start_ns = auxi
    if auxi_rate and rate and auxi_rate != rate:
        raise CaptureError(f"'{path}': the fmt chunk says {rate} samples/s and the auxi chunk "
                           f"{auxi_rate}; state the rate to settle which")
    if centre:
        out.center_hz = centre
    else:
        out.notes.append(f"'{path}': its auxi centre frequency is zero, which means unknown "
                         "(give center_hz=)")
    if start_ns is not None:
        out.anchor_ns = start_ns
        out.anchor_accuracy_ns = RECORDER_CLOCK_ACCURACY_NS
    return out


class WavFile:
    """A WAV recording as a sample source: memory-mapped blocks of its data."""

    def __init__(self, path: str, rate: int | None = None, center_hz: int | None = None):
        self.meta = m = read_wav_metadata(path)
        if rate is not None and m.rate is not None and rate != m.rate:
            raise CaptureError(f"rate={rate} contradicts the fmt chunk's {m.rate}")
        self.rate = m.rate if m.rate is not None else rate
        if self.rate is None:
            raise CaptureError(f"'{path}' states no sample rate: give one")
        if center_hz is not None and m.center_hz is not None and center_hz != m.center_hz:
            raise CaptureError(f"center={center_hz} contradicts the auxi chunk's {m.center_hz}")
        self.center_hz = m.center_hz if m.center_hz is not None else center_hz
        self.anchor_ns = m.anchor_ns
        self.format = m.format
        self.length = m.total_samples
        self._bps = BYTES_PER_SAMPLE[self.format]
        self._data = (np.memmap(path, dtype=np.uint8, mode="r", offset=m.data_offset,
                                shape=(self.length * self._bps,))
                      if self.length else np.zeros(0, np.uint8))

    def codes(self, start: int, count: int) -> np.ndarray:
        if start < 0 or count < 0 or start + count > self.length:
            raise CaptureError(f"samples [{start}, {start + count}) are outside the "
                               f"{self.length} the recording holds")
        a = start * self._bps
        return _view_codes(np.asarray(self._data[a:a + count * self._bps]), self.format)

    def blocks(self, block_samples: int, start: int = 0) -> Iterator[SourceBlock]:
        at = start
        while at < self.length:
            n = min(block_samples, self.length - at)
            yield SourceBlock(self.codes(at, n), at, n, self.center_hz)
            at += n


def write_wav(path: str, codes: np.ndarray, fmt: str, rate: int, center_hz: int = 0,
              start_ns: int | None = None, container: str = "RIFF",
              extensible: bool = False) -> None:
    """An IQ WAV (tests and tools): codes are the native samples' bytes in
    order; RF64/BW64 put the sizes in ds64; the auxi chunk carries the centre
    and the start time when either is given."""
    tag, bits = {"cu8": (TAG_PCM, 8), "cs16": (TAG_PCM, 16), "cs24": (TAG_PCM, 24),
                 "cf32": (TAG_FLOAT, 32)}[fmt]
    payload = np.ascontiguousarray(codes).view(np.uint8).tobytes()
    align = 2 * (bits // 8)
    if extensible:
        fmt_body = struct.pack("<HHIIHHHHI", TAG_EXTENSIBLE, 2, rate, rate * align, align, bits,
                               22, bits, 3) + struct.pack("<H", tag) + bytes(14)
    else:
        fmt_body = struct.pack("<HHIIHH", tag, 2, rate, rate * align, align, bits)
    chunks = [(b"fmt ", fmt_body)]
    if center_hz or start_ns is not None:
        st = bytes(16)
        if start_ns is not None:
            t = __import__("time").gmtime(start_ns // 1_000_000_000)
            st = struct.pack("<8H", t.tm_year, t.tm_mon, (t.tm_wday + 1) % 7, t.tm_mday,
                             t.tm_hour, t.tm_min, t.tm_sec, (start_ns // 1_000_000) % 1000)
        chunks.append((b"auxi", st + bytes(16) + struct.pack("<II", center_hz, rate) +
                       bytes(AUXI_CORE - 40)))
    wide = container != "RIFF"
    body = b""
    for cid, data in chunks:
        body += cid + struct.pack("<I", len(data)) + data + (b"\0" if len(data) & 1 else b"")
    data_size = len(payload)
    if wide:
        riff_size = 4 + 8 + DS64_CORE + len(body) + 8 + data_size + (data_size & 1)
        ds64 = struct.pack("<QQQI", riff_size, data_size, data_size // align, 0)
        body = b"ds64" + struct.pack("<I", DS64_CORE) + ds64 + body
        header = container.encode() + struct.pack("<I", SIZE_SENTINEL) + b"WAVE"
        data_hdr = b"data" + struct.pack("<I", SIZE_SENTINEL)
    else:
        riff_size = 4 + len(body) + 8 + data_size + (data_size & 1)
        header = b"RIFF" + struct.pack("<I", riff_size) + b"WAVE"
        data_hdr = b"data" + struct.pack("<I", data_size)
    with open(path, "wb") as f:
        f.write(header + body + data_hdr + payload + (b"\0" if data_size & 1 else b""))
