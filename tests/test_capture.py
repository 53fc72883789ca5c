"""The recordings around the path: the SigMF reader against the reference's
own cases (tests/engine/test_sources.cpp:803-1086), the rolling capture
writer and its replay (tests/capture/test_writer.cpp:154-665): segments roll
on retune, rate, size and loss with the reason recorded; a meta carries the
keys format.h names and the reader reads it; stems parse back; recovery
finalises an open segment; retention deletes the oldest first; a written
segment and a 24-bit capture replay sample for sample."""
import json
import os

import numpy as np
import pytest

from revenant_amd.capture import (CaptureError, CaptureReplay, CaptureWriter, SigmfFile,
                                  SourceBlock, catalog, format_sigmf_datetime, parse_sigmf_datetime,
                                  parse_stem, read_sigmf_metadata, segment_stem,
                                  sigmf_datatype_to_format, sigmf_meta_path_for, SegmentInfo)


def ramp_cf32(n, offset):
    v = np.arange(offset, offset + 2 * n, dtype=np.float32)
    return v.reshape(n, 2)


def write_pair(tmp_path, stem, samples, datatype, rate, captures, extra=None):
    data = tmp_path / f"{stem}.sigmf-data"
    meta = tmp_path / f"{stem}.sigmf-meta"
    data.write_bytes(np.ascontiguousarray(samples).tobytes())
    glob = {"core:datatype": datatype, "core:sample_rate": rate, "core:version": "1.0.0"}
    glob.update(extra or {})
    meta.write_text(json.dumps({"global": glob, "captures": captures, "annotations": []}))
    return str(data), str(meta)


def test_datatypes_map_to_the_convert_formats_and_the_rest_are_refused():
    assert [sigmf_datatype_to_format(d) for d in ("cu8", "ci8", "ci16_le", "ci24_le", "cf32_le",
                                                   "cu8_be", "ci8_le")] == \
        ["cu8", "cs8", "cs16", "cs24", "cf32", "cu8", "cs8"]
    for bad, word in (("cf32_be", "big-endian"), ("rf32_le", "real"), ("ci32_le", "no convert"),
                      ("x16", "does not parse"), ("ci16_xe", "does not parse")):
        with pytest.raises(CaptureError) as e:
            sigmf_datatype_to_format(bad)
        assert word in str(e.value) and bad in str(e.value)


def test_datetime_is_utc_with_a_z_and_round_trips():
    from datetime import datetime, timezone
    want = datetime(2026, 9, 21, 14, 3, 7, tzinfo=timezone.utc)
    assert parse_sigmf_datetime("2026-09-21T14:03:07.250Z") == \
        int(want.timestamp()) * 1_000_000_000 + 250_000_000
    assert parse_sigmf_datetime("1970-01-01T00:00:00Z") == 0
    ns = 1_790_000_587_123_456_789
    assert parse_sigmf_datetime(format_sigmf_datetime(ns)) == ns
    for bad in ("2026-09-21T14:03:07.250", "2026-09-21T14:03:07.Z", "2026-09-21 14:03:07Z",
                "2026-09-21T14:03:07+02:00"):
        with pytest.raises(CaptureError):
            parse_sigmf_datetime(bad)


def test_a_sigmf_pair_supplies_rate_format_centre_and_anchor(tmp_path):
    samples = ramp_cf32(20_000, 0)
    data, _ = write_pair(tmp_path, "capture", samples, "cf32_le", 2_000_000,
                         [{"core:sample_start": 0, "core:frequency": 7_074_000,
                           "core:datetime": "2026-09-21T14:03:07.250Z"}])
    f = SigmfFile(data)
    assert (f.rate, f.center_hz, f.format, f.length) == (2_000_000, 7_074_000, "cf32", 20_000)
    assert f.anchor_ns == parse_sigmf_datetime("2026-09-21T14:03:07.250Z")
    got = np.concatenate([b.codes for b in f.blocks(4096)]).reshape(-1, 2)
    assert np.array_equal(got, samples)


def test_a_dataset_with_no_sidecar_names_the_file_it_looked_for(tmp_path):
    data = tmp_path / "orphan.sigmf-data"
    data.write_bytes(ramp_cf32(1000, 0).tobytes())
    with pytest.raises(CaptureError) as e:
        SigmfFile(str(data))
    assert "orphan.sigmf-meta" in str(e.value) and "container=raw" in str(e.value)
    assert sigmf_meta_path_for("x.sigmf-data") == "x.sigmf-meta"
    assert sigmf_meta_path_for("x.bin") == "x.bin.sigmf-meta"


def test_a_sidecar_that_outruns_its_dataset_is_refused_with_both_lengths(tmp_path):
    data, meta = write_pair(tmp_path, "cut", ramp_cf32(5000, 0), "cf32_le", 2_000_000,
                            [{"core:sample_start": 0, "core:frequency": 14_074_000},
                             {"core:sample_start": 900_000, "core:frequency": 14_074_000}])
    with pytest.raises(CaptureError) as e:
        read_sigmf_metadata(meta)
    msg = str(e.value)
    assert "900000" in msg and "5000" in msg and "disagree" in msg


def test_a_rate_that_contradicts_the_sidecar_is_refused(tmp_path):
    data, _ = write_pair(tmp_path, "clash", ramp_cf32(4000, 0), "cf32_le", 2_000_000,
                         [{"core:sample_start": 0, "core:frequency": 10_000_000}])
    with pytest.raises(CaptureError) as e:
        SigmfFile(data, rate=2_400_000)
    assert "2000000" in str(e.value) and "2400000" in str(e.value)


def test_a_recording_that_retunes_is_refused_until_a_segment_is_named(tmp_path):
    samples = ramp_cf32(30_000, 0)
    data, _ = write_pair(tmp_path, "hop", samples, "cf32_le", 2_000_000,
                         [{"core:sample_start": 0, "core:frequency": 7_074_000},
                          {"core:sample_start": 10_000, "core:frequency": 14_074_000}])
    with pytest.raises(CaptureError) as e:
        SigmfFile(data)
    msg = str(e.value)
    for word in ("retunes", "segment=0", "segment=1", "7074000", "14074000"):
        assert word in msg
    f = SigmfFile(data, segment=1)
    assert f.center_hz == 14_074_000 and f.length == 20_000
    got = np.concatenate([b.codes for b in f.blocks(4096)]).reshape(-1, 2)
    assert np.array_equal(got, samples[10_000:])
    with pytest.raises(CaptureError) as e:
        SigmfFile(data, segment=7)
    assert "has 2" in str(e.value)


def test_several_segments_at_one_frequency_play_as_one_stream(tmp_path):
    samples = ramp_cf32(8000, 3)
    data, _ = write_pair(tmp_path, "split", samples, "cf32_le", 384_000,
                         [{"core:sample_start": 0, "core:frequency": 3_573_000},
                          {"core:sample_start": 4000, "core:frequency": 3_573_000}])
    f = SigmfFile(data)
    assert f.center_hz == 3_573_000 and f.length == 8000
    assert any("2 capture segments" in n for n in f.meta.notes)
    got = np.concatenate([b.codes for b in f.blocks(1024)])
    assert got[-1] == samples.ravel()[-1] and got.size == samples.size


def test_sidecar_checks(tmp_path):
    base = [{"core:sample_start": 0, "core:frequency": 1_000_000}]
    cases = [
        ({"core:version": "2.0.0"}, base, "version 1"),
        ({"core:metadata_only": True}, base, "metadata_only"),
        ({"core:num_channels": 2}, base, "num_channels"),
        ({"core:sample_rate": 2_000_000.5}, base, "whole number"),
        ({"core:trailing_bytes": 10**9}, base, "trailing"),
        ({}, [{"core:sample_start": 0, "core:header_bytes": 16}], "header_bytes"),
        ({}, [{"core:sample_start": 0, "core:frequency": 1e6 + 0.5}], "whole number of hertz"),
        ({}, [{"core:sample_start": 10}, {"core:sample_start": 5}], "sorted"),
        ({}, [{"core:frequency": 5}], "sample_start"),
        ({"core:offset": 100}, [{"core:sample_start": 50}], "core:offset"),
        ({}, [{"core:sample_start": 0, "core:datetime": "2026-01-01T00:00:00.5"}], "'Z'"),
    ]
    for k, (extra, caps, word) in enumerate(cases):
        _, meta = write_pair(tmp_path, f"c{k}", ramp_cf32(100, 0), "cf32_le", 1_000_000, caps,
                             extra)
        with pytest.raises(CaptureError) as e:
            read_sigmf_metadata(meta)
        assert word in str(e.value), (k, str(e.value))
    # trailing bytes and an offset are honoured; an empty captures array is a note
    _, meta = write_pair(tmp_path, "ok", ramp_cf32(100, 0), "cf32_le", 1_000_000,
                         [{"core:sample_start": 40}], {"core:trailing_bytes": 80,
                                                       "core:offset": 40})
    m = read_sigmf_metadata(meta)
    assert m.total_samples == 90 and m.segments[0].sample_start == 0
    _, meta = write_pair(tmp_path, "empty", ramp_cf32(100, 0), "cf32_le", 1_000_000, [])
    assert "captures is empty" in read_sigmf_metadata(meta).notes[0]


# ---------------------------------------------------------------------------
# The rolling capture and its replay
# ---------------------------------------------------------------------------

RATE = 2_400_000
ANCHOR = 1_790_000_000_000_000_000


def cu8_block(first, n, centre, seed, dropped=0):
    codes = np.random.default_rng(seed).integers(0, 256, 2 * n, dtype=np.uint8)
    return SourceBlock(codes, first, n, centre, dropped)


def test_segments_roll_on_retune_rate_size_and_loss(tmp_path):
    w = CaptureWriter(str(tmp_path), max_segment_bytes=2 * 8192 + 100, stream_epoch=0xABC)
    at = 0
    for k in range(3):
        w.write(cu8_block(at, 4096, 100_000_000, k), RATE, "cu8", ANCHOR)
        at += 4096
    w.write(cu8_block(at, 4096, 101_000_000, 9), RATE, "cu8", ANCHOR)          # retune
    at += 4096
    w.write(cu8_block(at, 4096, 101_000_000, 10), 2_048_000, "cu8", ANCHOR)   # rate
    at += 4096
    w.write(cu8_block(at + 500, 4096, 101_000_000, 11, dropped=500), 2_048_000, "cu8", ANCHOR)
    w.close("stop")
    segs = catalog(tmp_path)
    reasons = [json.loads(open(s.meta_path).read())["global"]["revenant:close_reason"]
               for s in segs]
    assert reasons == ["size", "retune", "rate", "loss", "stop"]
    assert [s.info.sample_count for s in segs] == [8192, 4096, 4096, 4096, 4096]
    assert segs[-1].info.dropped_before == 500
    assert all(s.info.stream_epoch == 0xABC for s in segs)
    assert not list(tmp_path.glob("*/*.open*"))


def test_a_meta_carries_every_key_and_the_reader_reads_it(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=7, hw="rtl-sdr v4")
    w.write(cu8_block(1000, 2048, 97_300_000, 1), RATE, "cu8", ANCHOR)
    path = w.close("stop")
    meta = read_sigmf_metadata(sigmf_meta_path_for(str(path)))
    g = meta.global_
    for key in ("core:datatype", "core:sample_rate", "core:version", "core:recorder", "core:hw",
                "core:extensions", "revenant:stream_epoch", "revenant:first_index",
                "revenant:epoch_anchor_ns", "revenant:dropped_before",
                "revenant:close_reason"):
        assert key in g, key
    assert g["core:extensions"][0]["name"] == "revenant" and g["core:extensions"][0]["optional"]
    assert (meta.rate, meta.center_hz, meta.format, meta.total_samples) == \
        (RATE, 97_300_000, "cu8", 2048)
    info = SegmentInfo(7, 1000, 2048, ANCHOR, RATE, 97_300_000, "cu8")
    assert meta.anchor_ns == info.start_ns()
    assert path.parent.name == format_sigmf_datetime(info.start_ns())[:10]


def test_stems_parse_back_to_what_made_them():
    info = SegmentInfo(0x0123456789ABCDEF, 123_456_789, 0, ANCHOR, RATE, 145_800_000, "cs16")
    ns, centre, epoch, first = parse_stem(segment_stem(info))
    assert (centre, epoch, first) == (145_800_000, 0x0123456789ABCDEF, 123_456_789)
    assert ns == info.start_ns() // 1_000_000 * 1_000_000


def test_recovery_finalises_a_segment_a_dead_writer_left_open(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=5)
    w.write(cu8_block(0, 3000, 50_000_000, 3), RATE, "cu8", ANCHOR)
    w._file.write(b"\x01")  # a torn half sample
    w._file.flush()
    w._file = None  # the process dies here: no close
    w2 = CaptureWriter(str(tmp_path), stream_epoch=6)
    assert len(w2.recovered) == 1
    segs = catalog(tmp_path)
    assert len(segs) == 1 and segs[0].info.sample_count == 3000
    g = json.loads(open(segs[0].meta_path).read())["global"]
    assert g["revenant:close_reason"] == "recovered"
    assert not list(tmp_path.glob("*/*.open*"))


def test_retention_deletes_the_oldest_closed_segments_first(tmp_path):
    # room for three segments with their metas
    w = CaptureWriter(str(tmp_path), max_segment_bytes=8192, retention_bytes=3 * (8192 + 2048),
                      stream_epoch=1)
    for k in range(6):
        w.write(cu8_block(k * 4096, 4096, 10_000_000, k), RATE, "cu8", ANCHOR)
    w.close("stop")
    segs = catalog(tmp_path)
    assert [s.info.first_index for s in segs] == [3 * 4096, 4 * 4096, 5 * 4096]
    assert w.segments_deleted == 3


def test_a_written_segment_replays_through_the_file_reader_sample_for_sample(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=2)
    blocks = [cu8_block(k * 1000, 1000, 7_000_000, 20 + k) for k in range(5)]
    for b in blocks:
        w.write(b, RATE, "cu8", ANCHOR)
    path = w.close("stop")
    f = SigmfFile(str(path))
    got = np.concatenate([b.codes for b in f.blocks(777)])
    assert np.array_equal(got, np.concatenate([b.codes for b in blocks]))


def test_a_24_bit_capture_replays_through_capture_replay_sample_for_sample(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=3)
    rng = np.random.default_rng(4)
    sent = []
    for k in range(4):
        raw = rng.integers(0, 256, 6 * 1500, dtype=np.uint8)
        sent.append(raw)
        w.write(SourceBlock(raw, k * 1500, 1500, 14_200_000), 192_000, "cs24", ANCHOR)
    w.close("stop")
    r = CaptureReplay(str(tmp_path))
    assert r.format == "cs24" and r.rate == 192_000
    got = np.concatenate([b.codes for b in r.blocks(1024)])
    assert np.array_equal(got, np.concatenate(sent))


def test_replay_follows_a_retune_reports_a_gap_and_stops_at_a_rate_change(tmp_path):
    w = CaptureWriter(str(tmp_path), stream_epoch=9)
    w.write(cu8_block(0, 2000, 1_000_000, 1), RATE, "cu8", ANCHOR)
    w.write(cu8_block(2000, 2000, 2_000_000, 2), RATE, "cu8", ANCHOR)            # retune
    w.write(cu8_block(4700, 2000, 2_000_000, 3, dropped=700), RATE, "cu8", ANCHOR)  # gap
    w.write(cu8_block(6700, 2000, 2_000_000, 4), 1_200_000, "cu8", ANCHOR)       # rate
    w.close("stop")
    r = CaptureReplay(str(tmp_path))
    blocks = list(r.blocks(1500))
    assert [b.first_index for b in blocks] == [0, 1500, 2000, 3500, 4700, 6200]
    assert [b.center_hz for b in blocks] == [1_000_000] * 2 + [2_000_000] * 4
    assert [b.dropped_before for b in blocks] == [0, 0, 0, 0, 700, 0]
    assert any("rate change" in n for n in r.notes)
    # from= lands inside a segment, on the sample at that instant
    info0 = catalog(tmp_path)[0].info
    start = info0.start_ns() + 1000 * 1_000_000_000 // RATE
    r2 = CaptureReplay(str(tmp_path), from_ns=start)
    assert r2.first_index == 1000
    assert next(iter(r2.blocks(100))).first_index == 1000
