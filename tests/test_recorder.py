"""The queued capture recorder (core/capture/writer.h SegmentWriter): the
cases of tests/capture/test_writer.cpp:154-440 run against
revenant_amd.recorder.SegmentRecorder. CPU only: files under a temporary root.
"""
import json
import threading
import time
from pathlib import Path

import numpy as np
import pytest

from revenant_amd.capture import (CaptureWriter, SourceBlock, catalog, parse_stem,
                                  read_sigmf_metadata)
from revenant_amd.recorder import SegmentRecorder
from revenant_amd.search import DETECT_GENERATOR, Annotation, read_annotations

RATE = 1_000_000
ANCHOR = 1_791_806_400_000_000_000   # 2026-10-12T12:00:00Z
BLOCK = 1000
HW = 'test "dongle"\nline two'


def block_codes(start: int, count: int) -> np.ndarray:
    """Each sample's two bytes are its index's low bytes, so a sample out of
    place or a repeated block shows (test_writer.cpp block_bytes)."""
    idx = np.arange(start, start + count, dtype=np.int64)
    out = np.empty(2 * count, dtype=np.uint8)
    out[0::2], out[1::2] = idx & 0xFF, (idx >> 8) & 0xFF
    return out


def feed(rec, start, centre, dropped=0, rate=RATE, anchor=ANCHOR):
    rec.consume(SourceBlock(block_codes(start, BLOCK), start, BLOCK, centre, dropped), rate, "cu8",
                anchor)


def segments(root):
    """(meta path, global, doc) of every closed segment, in stream order."""
    out = []
    for meta in Path(root).glob("*/*.sigmf-meta"):
        doc = json.loads(meta.read_text())
        out.append((meta, doc["global"], doc))
    out.sort(key=lambda s: parse_stem(s[0].name[:-len(".sigmf-meta")])[3])
    return out


def wait_for(rec, done, seconds=5.0):
    end = time.monotonic() + seconds
    while time.monotonic() < end:
        if done(rec.stats()):
            return True
        time.sleep(0.001)
    return done(rec.stats())


def test_segments_roll_on_retune_rate_size_and_loss_with_the_reason_recorded(tmp_path):
    rec = SegmentRecorder(str(tmp_path), max_segment_bytes=4 * BLOCK * 2, stream_epoch=0x1234,
                          hw=HW)
    rec.start()
    at = 0
    for _ in range(5):               # four fill a segment, the fifth rolls it
        feed(rec, at, 100_000_000)
        at += BLOCK
    feed(rec, at, 101_000_000)       # retune
    at += BLOCK
    feed(rec, at + 500, 101_000_000, dropped=500)   # the device lost 500 samples
    at += 500 + BLOCK
    feed(rec, at, 101_000_000, rate=2_000_000)      # rate change
    rec.stop()
    rec.stop()                       # safe twice
    st = rec.stats()
    assert st.last_error == "" and st.segments_written == 5 and st.blocks_dropped == 0
    segs = segments(tmp_path)
    assert [g["revenant:close_reason"] for _, g, _ in segs] == \
        ["size", "retune", "loss", "rate", "stop"]
    assert segs[3][1]["revenant:dropped_before"] == 500
    assert segs[1][1]["revenant:dropped_before"] == 0
    assert not [p for p in tmp_path.rglob("*") if ".open" in p.name]


def test_a_meta_carries_every_key_and_the_reader_reads_it(tmp_path):
    rec = SegmentRecorder(str(tmp_path), stream_epoch=0x1234, hw=HW)
    rec.set_gain_db(29.7)
    rec.start()
    for at in range(0, 4 * BLOCK, BLOCK):
        feed(rec, at, 145_500_000)
    rec.add_annotation(2500, Annotation(sample_count=1500, freq_lower=145_490_000,
                                        freq_upper=145_510_000, confidence=0.75, peak_db=-20.5,
                                        snr_db=18.0, family="nbfm"), track_id=42)
    rec.stop()
    [(meta, g, doc)] = segments(tmp_path)
    assert g["core:datatype"] == "cu8" and g["core:sample_rate"] == RATE
    assert g["core:version"].startswith("1.") and g["core:recorder"].startswith("Revenant ")
    assert g["core:hw"] == HW
    ext = g["core:extensions"][0]
    assert (ext["name"], ext["version"], ext["optional"]) == ("revenant", "1.1.0", True)
    assert (g["revenant:stream_epoch"], g["revenant:first_index"]) == (0x1234, 0)
    assert g["revenant:epoch_anchor_ns"] == ANCHOR and g["revenant:dropped_before"] == 0
    assert g["revenant:close_reason"] == "stop" and g["revenant:gain_db"] == 29.7
    cap = doc["captures"][0]
    assert cap["core:frequency"] == 145_500_000
    assert cap["core:datetime"].startswith("2026-10-12T12:00:00")
    [a] = doc["annotations"]
    assert (a["core:sample_start"], a["core:sample_count"]) == (2500, 1500)
    assert a["core:freq_lower_edge"] == 145_490_000 and a["core:label"] == "nbfm"
    assert a["core:generator"] == DETECT_GENERATOR and a["revenant:confidence"] == 0.75
    assert a["revenant:family"] == "nbfm" and a["revenant:track_id"] == 42
    # the stem follows the format and sits in its UTC day directory
    assert meta.parent.name == "2026-10-12"
    assert meta.name == "rev-20261012T120000.000Z-145500000-0000000000001234-0.sigmf-meta"
    m = read_sigmf_metadata(str(meta))
    assert (m.rate, m.format, m.center_hz, m.anchor_ns) == (RATE, "cu8", 145_500_000, ANCHOR)
    data = meta.with_name(meta.name.replace(".sigmf-meta", ".sigmf-data"))
    assert m.data_bytes == data.stat().st_size == 4 * BLOCK * 2


def test_a_track_spanning_closed_segments_is_filed_in_every_one_with_its_id(tmp_path):
    rec = SegmentRecorder(str(tmp_path), max_segment_bytes=BLOCK * 2, stream_epoch=0x1234)
    rec.start()
    for at in range(0, 4 * BLOCK, BLOCK):
        feed(rec, at, 98_100_000)
    # three closed and the fourth open before the track is reported
    assert wait_for(rec, lambda s: s.segments_written == 3)
    wfm = Annotation(sample_count=3000, freq_lower=98_000_000, freq_upper=98_200_000,
                     confidence=0.9, family="wfm")
    rec.add_annotation(500, wfm, 7)
    assert wait_for(rec, lambda s: s.annotations_merged_late == 3)
    # the first segment goes; a second track reaching back into it loses that
    # part and keeps the rest
    segs = segments(tmp_path)
    assert len(segs) == 3
    first = segs[0][0]
    first.unlink()
    first.with_name(first.name.replace(".sigmf-meta", ".sigmf-data")).unlink()
    rec.add_annotation(900, wfm, 8)
    rec.stop()
    st = rec.stats()
    assert st.annotations_discarded == 1 and st.annotations_merged_late == 3 + 2
    segs = segments(tmp_path)
    assert len(segs) == 3
    seven = {1: (0, 1000), 2: (0, 1000), 3: (0, 500)}
    eight = {1: (0, 1000), 2: (0, 1000), 3: (0, 900)}
    for i, (meta, _, _) in enumerate(segs, start=1):
        stored = read_annotations(str(meta))
        assert len(stored) == 2
        for s in stored:
            assert s.generator == DETECT_GENERATOR and s.annotation.family == "wfm"
            assert s.track_id in (7, 8)
            want = (seven if s.track_id == 7 else eight)[i]
            assert (s.annotation.sample_start, s.annotation.sample_count) == want
        read_sigmf_metadata(str(meta))  # a late append keeps the meta whole


def test_a_full_queue_drops_counts_and_closes_the_segment_as_loss(tmp_path, monkeypatch):
    rec = SegmentRecorder(str(tmp_path), queue_blocks=1, stream_epoch=0x1234)
    rec.start()
    # The recorder thread is held inside its first write, so the one-block
    # queue fills and every block after that is dropped.
    entered, release = threading.Event(), threading.Event()
    real = rec._write

    def held(item):
        if not entered.is_set():
            entered.set()
            release.wait(10)
        real(item)

    monkeypatch.setattr(rec, "_write", held)
    feed(rec, 0, 50_000_000)
    assert entered.wait(5)
    at = BLOCK
    for _ in range(10):
        feed(rec, at, 50_000_000)
        at += BLOCK
    release.set()
    assert wait_for(rec, lambda s: s.bytes_written == 2 * BLOCK * 2)
    feed(rec, at, 50_000_000)   # once drained: the loss is attributed to a segment
    rec.stop()
    st = rec.stats()
    assert st.blocks_dropped == 9 and st.samples_dropped == 9 * BLOCK
    assert st.bytes_written + st.samples_dropped * 2 == 12 * BLOCK * 2
    segs = segments(tmp_path)
    assert sum(g["revenant:close_reason"] == "loss" for _, g, _ in segs) >= 1
    assert sum(g["revenant:dropped_before"] for _, g, _ in segs) == st.samples_dropped
    for meta, _, _ in segs:
        m = read_sigmf_metadata(str(meta))
        data = meta.with_name(meta.name.replace(".sigmf-meta", ".sigmf-data"))
        assert m.data_bytes == data.stat().st_size
    # the segment after the gap starts where the dropped blocks ended
    assert segs[-1][1]["revenant:first_index"] == 11 * BLOCK


def test_retention_deletes_the_oldest_closed_segments_first(tmp_path):
    budget = 3 * (BLOCK * 2 + 1500)   # room for about three segments with their metas
    rec = SegmentRecorder(str(tmp_path), max_segment_bytes=BLOCK * 2, retention_bytes=budget,
                          stream_epoch=0x1234)
    rec.start()
    for at in range(0, 10 * BLOCK, BLOCK):
        feed(rec, at, 433_920_000)
    rec.stop()
    st = rec.stats()
    assert st.segments_written == 10 and st.segments_deleted >= 6
    assert st.retained_bytes <= budget
    segs = segments(tmp_path)
    assert len(segs) == 10 - st.segments_deleted
    assert segs[-1][1]["revenant:first_index"] == 9000
    assert segs[0][1]["revenant:first_index"] != 0


def test_retention_by_age_removes_old_segments_and_their_empty_day_directory(tmp_path):
    old = SegmentRecorder(str(tmp_path), stream_epoch=0x1234)
    old.start()
    feed(old, 0, 7_100_000, anchor=ANCHOR - 3 * 86_400 * 10**9)
    old.stop()
    assert (tmp_path / "2026-10-09").exists()
    rec = SegmentRecorder(str(tmp_path), stream_epoch=0x5678, retention_max_age_ns=86_400 * 10**9,
                          now_ns=lambda: ANCHOR + 10**9)
    rec.start()
    feed(rec, 0, 7_100_000)
    rec.stop()
    assert rec.stats().segments_deleted == 1
    assert not (tmp_path / "2026-10-09").exists()
    assert len(segments(tmp_path)) == 1


def test_a_write_error_stops_writing_and_the_stream_carries_on(tmp_path, monkeypatch):
    rec = SegmentRecorder(str(tmp_path), stream_epoch=0x1234)
    rec.start()
    feed(rec, 0, 7_000_000)
    assert wait_for(rec, lambda s: s.bytes_written == BLOCK * 2)

    def full_disk(self, *args, **kwargs):
        raise OSError(28, "No space left on device")

    monkeypatch.setattr(CaptureWriter, "write", full_disk)
    feed(rec, BLOCK, 7_000_000)
    assert wait_for(rec, lambda s: s.last_error != "")
    for k in range(2, 5):  # offered after the error: counted, never raised
        feed(rec, k * BLOCK, 7_000_000)
    rec.stop()
    st = rec.stats()
    assert "No space left" in st.last_error and st.blocks_unwritten == 4
    assert st.bytes_written == BLOCK * 2


def test_recovery_at_start_keeps_the_gain_and_the_hardware_string(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=5, hw=HW)
    w.gain_db = 12.5
    w.write(SourceBlock(block_codes(0, 3000), 0, 3000, 50_000_000), RATE, "cu8", ANCHOR)
    w._file.flush()
    w._file = None  # the process dies with the segment open
    rec = SegmentRecorder(str(tmp_path), stream_epoch=6)
    rec.start()
    rec.stop()
    [seg] = catalog(tmp_path)
    g = json.loads(Path(seg.meta_path).read_text())["global"]
    assert g["revenant:close_reason"] == "recovered" and seg.info.sample_count == 3000
    assert g["core:hw"] == HW and g["revenant:gain_db"] == 12.5


def test_a_segment_closes_at_its_duration_limit(tmp_path):
    rec = SegmentRecorder(str(tmp_path), max_segment_seconds=0.0025, stream_epoch=1)
    rec.start()
    for at in range(0, 6 * BLOCK, BLOCK):
        feed(rec, at, 1_000_000)
    rec.stop()
    segs = segments(tmp_path)
    # 2500 samples at 1 MS/s: a segment takes two blocks and closes as "size"
    assert [g["revenant:close_reason"] for _, g, _ in segs] == ["size", "size", "stop"]
    assert [g["revenant:first_index"] for _, g, _ in segs] == [0, 2000, 4000]


def test_a_zero_block_queue_is_refused():
    from revenant_amd.capture import CaptureError

    with pytest.raises(CaptureError):
        SegmentRecorder("/tmp/unused", queue_blocks=0)
