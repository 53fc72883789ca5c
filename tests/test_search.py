"""Annotations of closed segments and the questions asked of them
(revenant_amd/search.py), after tests/capture/test_search.cpp:138-700: the
query and occupancy cases over hand-made annotations (so each asserts
arithmetic, not a detector), and the merge cases holding the writer to
leaving what it does not own alone."""
import json
import math

import numpy as np
import pytest

from revenant_amd.capture import CaptureError, CaptureWriter, SegmentInfo, SourceBlock
from revenant_amd.search import (DECODE_GENERATOR, DETECT_GENERATOR, INDEX_GENERATOR, Annotation,
                                 IndexedSegment, IndexStamp, OccupancyRequest, Query,
                                 StoredAnnotation, append_annotations, append_annotations_to,
                                 indexed_by, indexed_segments, merge_annotations, occupancy,
                                 read_annotations, read_index_stamps, run_query,
                                 write_annotations)

SECOND = 1_000_000_000
ANCHOR = 1_760_227_200 * SECOND  # 2026-10-12T00:00:00Z


def _info(first, count, rate=1_000_000):
    return SegmentInfo(stream_epoch=0x1234, first_index=first, sample_count=count,
                       epoch_anchor_ns=ANCHOR, rate=rate, center=146_000_000, format="cf32")


def _stored(start, count, low, high, conf, snr, family="", gen=INDEX_GENERATOR):
    return StoredAnnotation(Annotation(start, count, low, high, conf, -40.0, snr, family), gen)


def _two_segments():
    """Two ten-second segments at 1 MS/s back to back in one epoch: every
    annotation's wall clock is the anchor plus its stream second."""
    return [
        IndexedSegment("a.sigmf-data", _info(0, 10_000_000), [
            _stored(1_000_000, 2_000_000, 146_010_000, 146_030_000, 0.9, 20.0, "NFM"),
            _stored(5_000_000, 500_000, 146_500_000, 146_512_500, 0.6, 8.0, "P25 voice"),
            _stored(1_000_000, 2_000_000, 146_010_000, 146_030_000, 0.95, 21.0, "NFM",
                    DETECT_GENERATOR)]),
        IndexedSegment("b.sigmf-data", _info(10_000_000, 10_000_000), [
            _stored(0, 4_000_000, 146_010_000, 146_030_000, 0.3, 4.0),
            _stored(2_000_000, 1_000_000, 145_000_000, 145_100_000, 0.99, 30.0, "DMR")]),
    ]


def test_an_empty_query_returns_every_annotation_in_time_order():
    hits = run_query(_two_segments(), Query())
    assert len(hits) == 5
    assert all(a.start_ns <= b.start_ns for a, b in zip(hits, hits[1:]))
    assert hits[0].start_ns == ANCHOR + SECOND and hits[-1].start_ns == ANCHOR + 12 * SECOND
    last = hits[-1]
    assert (last.data_path, last.end_ns, last.sample_start, last.first_index,
            last.stream_epoch) == ("b.sigmf-data", ANCHOR + 13 * SECOND, 2_000_000, 10_000_000,
                                   0x1234)


def test_time_and_frequency_are_overlap_tests():
    segs = _two_segments()
    hits = run_query(segs, Query(from_ns=ANCHOR + 2_500_000_000, to_ns=ANCHOR + 2_600_000_000))
    assert len(hits) == 2 and hits[0].family == "NFM"
    # ending exactly at from_ns, or starting exactly at to_ns, is outside
    assert run_query(segs, Query(from_ns=ANCHOR + 3 * SECOND, to_ns=ANCHOR + 5 * SECOND)) == []
    hits = run_query(segs, Query(freq_low=145_100_000, freq_high=145_200_000))  # touches an edge
    assert [h.family for h in hits] == ["DMR"]


def test_measure_filters_exclude_what_does_not_pass():
    segs = _two_segments()
    assert len(run_query(segs, Query(min_confidence=0.5))) == 4
    assert len(run_query(segs, Query(min_snr_db=10.0))) == 3
    assert len(run_query(segs, Query(min_duration_ns=2 * SECOND))) == 3
    assert [h.family for h in run_query(segs, Query(family="p25"))] == ["P25 voice"]
    hits = run_query(segs, Query(generator=DETECT_GENERATOR))
    assert len(hits) == 1 and hits[0].confidence == 0.95


def test_occupancy_counts_overlapping_presence_once_and_reports_coverage():
    segs = _two_segments()
    r = OccupancyRequest(ANCHOR, ANCHOR + 25 * SECOND, 5 * SECOND, 146_000_000, 146_040_000)
    b = occupancy(segs, r)
    assert len(b) == 5
    assert b[0].occupied == pytest.approx(0.4) and b[0].hits == 2 and b[0].captured == 1.0
    assert b[1].occupied == 0.0 and b[1].captured == 1.0
    assert b[2].occupied == pytest.approx(0.8)
    assert b[4].occupied == 0.0 and b[4].captured == 0.0  # nothing recorded: not "quiet"
    r.filter.min_confidence = 0.5
    assert occupancy(segs, r)[2].occupied == 0.0
    r.bucket_ns = 10 * SECOND
    b = occupancy(segs, r)
    assert len(b) == 3 and b[-1].end_ns == r.to_ns
    r.bucket_ns = 0
    with pytest.raises(CaptureError):
        occupancy(segs, r)
    with pytest.raises(CaptureError):
        occupancy(segs, OccupancyRequest(ANCHOR, ANCHOR, SECOND, 0, 1))


META = json.dumps({
    "global": {"core:datatype": "cf32_le", "core:sample_rate": 2400000.0,
               "core:version": "1.0.0", "revenant:stream_epoch": 4660,
               "someone_else:weird": {"nested": [1, 2.5, "a \"quoted\" é"]}},
    "captures": [{"core:sample_start": 0, "core:frequency": 146000000,
                  "core:datetime": "2026-10-12T00:00:00.000Z"}],
    "annotations": [
        {"core:sample_start": 500, "core:sample_count": 10, "core:freq_lower_edge": 1.0e8,
         "core:freq_upper_edge": 100010000, "core:generator": DETECT_GENERATOR,
         "revenant:track_id": 7, "revenant:confidence": 0.75},
        {"core:sample_start": 100, "core:sample_count": 10, "core:freq_lower_edge": 1,
         "core:freq_upper_edge": 2, "core:generator": INDEX_GENERATOR},
        {"core:sample_start": 900, "core:comment": "a person wrote this"}]})


def test_reindexing_replaces_only_the_index_entries(tmp_path):
    added = Annotation(700, 50, 146_010_000, 146_030_000, 0.875, -31.5, 18.25, "NFM")
    merged = merge_annotations(META, [added], INDEX_GENERATOR)
    doc, orig = json.loads(merged), json.loads(META)
    assert doc["global"] == orig["global"] and doc["captures"] == orig["captures"]
    assert {"core:sample_start": 900, "core:comment": "a person wrote this"} in doc["annotations"]
    path = tmp_path / "m.sigmf-meta"
    path.write_text(merged)
    read = read_annotations(str(path))
    assert [r.generator for r in read] == [DETECT_GENERATOR, INDEX_GENERATOR, ""]
    d, i, person = read
    assert (d.track_id, d.annotation.freq_lower, d.annotation.confidence) == (7, 100_000_000, 0.75)
    assert i.annotation == added
    assert person.annotation.sample_start == 900
    assert merge_annotations(merged, [added], INDEX_GENERATOR) == merged  # idempotent


def test_a_bare_meta_gets_an_array_and_a_stamp(tmp_path):
    path = tmp_path / "bare.sigmf-meta"
    path.write_text(json.dumps({"global": {"core:datatype": "cu8", "x:y": 1.5}, "captures": []}))
    write_annotations(str(path), [Annotation(3, 4, 5, 6, 0.5, 0.0, math.nan)])
    assert list(tmp_path.iterdir()) == [path]  # no temporary left behind
    doc = json.loads(path.read_text())
    assert doc["global"]["x:y"] == 1.5 and doc["global"]["revenant:indexed"][0]["generator"] == \
        INDEX_GENERATOR
    e = doc["annotations"][0]
    assert "revenant:snr_db" not in e and e["core:label"] == "signal"  # NaN is not written
    assert read_annotations(str(path))[0].annotation.sample_count == 4
    write_annotations(str(path), [])
    assert read_annotations(str(path)) == []


def test_a_pass_that_finds_nothing_is_recorded_as_having_run(tmp_path):
    first = IndexStamp(INDEX_GENERATOR, "2026-10-12T10:00:00.000Z", "channels=64")
    merged = merge_annotations(META, [], INDEX_GENERATOR, first)
    path = tmp_path / "s.sigmf-meta"
    path.write_text(merged)
    stamps, read = read_index_stamps(str(path)), read_annotations(str(path))
    assert stamps == [first]
    assert not any(r.generator == INDEX_GENERATOR for r in read)
    assert indexed_by(stamps, read) and not indexed_by([], read)
    other = merge_annotations(merged, [], "someone-else", IndexStamp("someone-else", "t", "s"))
    second = IndexStamp(INDEX_GENERATOR, "2026-10-13T00:00:00.000Z", "channels=32")
    again = merge_annotations(other, [], INDEX_GENERATOR, second)
    path.write_text(again)
    assert read_index_stamps(str(path)) == [IndexStamp("someone-else", "t", "s"), second]
    assert merge_annotations(again, [], INDEX_GENERATOR, second) == again


def test_write_annotations_stamps_a_quiet_pass(tmp_path):
    path = tmp_path / "q.sigmf-meta"
    path.write_text(META)
    write_annotations(str(path), [], INDEX_GENERATOR, "pace=0")
    (s,) = read_index_stamps(str(path))
    assert s.settings == "pace=0" and s.at.endswith("Z") and len(s.at) == 24


def test_a_malformed_meta_is_refused_and_left_untouched(tmp_path):
    path = tmp_path / "bad.sigmf-meta"
    bad = '{"global": {"core:datatype": "cu8"}, "annotations": ['
    path.write_text(bad)
    with pytest.raises(CaptureError):
        write_annotations(str(path), [])
    assert path.read_text() == bad
    with pytest.raises(CaptureError):
        read_annotations(str(path))
    path.write_text(json.dumps({"global": {}, "annotations": [{"core:sample_count": 3}]}))
    with pytest.raises(CaptureError, match="core:sample_start"):
        read_annotations(str(path))


def _decoded(start, decoder, fields):
    return StoredAnnotation(Annotation(start, 1_000_000, 852_000_000, 852_012_500, 1.0,
                                       decoder=decoder, decoded=fields, messages=4),
                            DECODE_GENERATOR)


def test_decoder_and_decoded_filters_match_whole_terms_anded_ignoring_case():
    segs = _two_segments()
    segs[0].annotations += [_decoded(2_000_000, "ais", [("mmsi", "366123456")]),
                            _decoded(2_000_000, "ais", [("callsign", "WDC1234")])]
    segs[1].annotations += [_decoded(1_000_000, "p25p1", [("talkgroup", "101"), ("source", "4500")]),
                            _decoded(3_000_000, "p25p1", [("talkgroup", "1011")])]
    hits = run_query(segs, Query(decoder="P25P1"))
    assert len(hits) == 2 and hits[0].decoder == "p25p1" and hits[0].generator == DECODE_GENERATOR
    hits = run_query(segs, Query(decoded=[("CallSign", "wdc1234")]))
    assert len(hits) == 1 and hits[0].decoded[0][1] == "WDC1234" and hits[0].messages == 4
    hits = run_query(segs, Query(decoded=[("talkgroup", "101")]))  # 101 does not find 1011
    assert [h.data_path for h in hits] == ["b.sigmf-data"]
    assert len(run_query(segs, Query(decoded=[("talkgroup", "101"), ("source", "4500")]))) == 1
    assert run_query(segs, Query(decoded=[("talkgroup", "101"), ("source", "4501")])) == []
    assert run_query(segs, Query(decoder="dmr", decoded=[("talkgroup", "101")])) == []
    assert len(run_query(segs, Query(generator=DETECT_GENERATOR))) == 1


def test_reindexing_leaves_decode_entries_and_appends_add(tmp_path):
    with_decode = append_annotations(META, [_decoded(300, "dmr", [("talkgroup", "9")])])
    merged = merge_annotations(with_decode, [Annotation(700, family="NFM")], INDEX_GENERATOR,
                               IndexStamp(INDEX_GENERATOR, "now"))
    path = tmp_path / "d.sigmf-meta"
    path.write_text(merged)
    read = read_annotations(str(path))
    assert sum(1 for r in read if r.generator == DECODE_GENERATOR and r.annotation.decoder ==
               "dmr" and r.annotation.decoded == [("talkgroup", "9")]) == 1
    append_annotations_to(str(path), [StoredAnnotation(Annotation(50, 5), DETECT_GENERATOR, 11)])
    read = read_annotations(str(path))
    assert read[0].track_id == 11 and len(read) == 5  # detect, new detect, decode, index, person
    assert [r.annotation.sample_start for r in read] == sorted(r.annotation.sample_start
                                                               for r in read)


def test_indexed_segments_read_a_capture_root(tmp_path):
    w = CaptureWriter(str(tmp_path), max_segment_bytes=1 << 20)
    codes = np.zeros(2 * 4096, dtype=np.uint8)
    for k in range(3):
        w.write(SourceBlock(codes, k * 4096, 4096, 100_000_000), 2_400_000, "cu8",
                epoch_anchor_ns=ANCHOR)
    w.close()
    (seg,) = indexed_segments(tmp_path)
    assert seg.info.sample_count == 3 * 4096 and seg.annotations == [] and seg.stamps == []
    meta = seg.data_path.replace(".sigmf-data", ".sigmf-meta")
    write_annotations(meta, [Annotation(100, 200, 100_000_000, 100_012_500, 0.9)])
    (seg,) = indexed_segments(tmp_path)
    assert indexed_by(seg.stamps, seg.annotations) and len(seg.annotations) == 1
    hits = run_query([seg], Query(freq_low=100_010_000, freq_high=100_011_000))
    assert len(hits) == 1 and hits[0].start_ns == ANCHOR + 100 * SECOND // 2_400_000
