"""Annotations of closed capture segments, and the questions asked of them
(core/capture/search.h, search.cpp; the meta layout of format.h:71-125).

A segment's .sigmf-meta carries core:annotations (each positioned by
core:sample_start / core:sample_count relative to the segment, with its
frequency edges, core:label, core:generator and the revenant keys: confidence,
peak_db, snr_db, family, track_id, decoder, decoded, messages) and, in its
global object, revenant:indexed: one stamp per offline pass ({generator, at,
settings}), so a quiet segment that was looked at differs from one nobody
looked at.

Writing keeps everything it does not own: other generators' entries, every
other key and value, through a temporary file renamed over the meta. Queries
are overlap tests in time and frequency; occupancy cuts a time range into
buckets and reports the share of each with something present, and the share
that was being recorded at all.
"""
from __future__ import annotations

import json
import math
import os
import tempfile
import time
from dataclasses import dataclass, field
from pathlib import Path

from .capture import CaptureError, SegmentInfo, catalog

DETECT_GENERATOR = "revenant-detect"
INDEX_GENERATOR = "revenant-index"
DECODE_GENERATOR = "revenant-decode"
INDEXED_KEY = "revenant:indexed"


@dataclass
class Annotation:
    """format.h:156 Annotation: relative to the segment's first sample."""
    sample_start: int = 0
    sample_count: int = 0
    freq_lower: int = 0
    freq_upper: int = 0
    confidence: float = 0.0
    peak_db: float = 0.0
    snr_db: float = 0.0
    family: str = ""
    decoder: str = ""
    decoded: list[tuple[str, str]] = field(default_factory=list)
    messages: int = 0


def annotation_label(a: Annotation) -> str:
    """core:label: the decode's name and identifiers, else the family, else 'signal'."""
    if a.decoder:
        return " ".join([a.decoder] + [f"{k} {v}" for k, v in a.decoded])
    return a.family or "signal"


@dataclass
class StoredAnnotation:
    annotation: Annotation
    generator: str = ""
    track_id: int = 0


@dataclass
class IndexStamp:
    generator: str
    at: str = ""
    settings: str = ""


@dataclass
class IndexedSegment:
    data_path: str
    info: SegmentInfo
    annotations: list[StoredAnnotation] = field(default_factory=list)
    stamps: list[IndexStamp] = field(default_factory=list)


def _entry(a: Annotation, generator: str, track_id: int = 0) -> dict:
    e = {"core:sample_start": int(a.sample_start), "core:sample_count": int(a.sample_count),
         "core:freq_lower_edge": int(a.freq_lower), "core:freq_upper_edge": int(a.freq_upper),
         "core:label": annotation_label(a), "core:generator": generator}
    for key, v in (("revenant:confidence", a.confidence), ("revenant:peak_db", a.peak_db),
                   ("revenant:snr_db", a.snr_db)):
        if math.isfinite(v):  # a measure that is not a number is left out
            e[key] = float(v)
    if a.family:
        e["revenant:family"] = a.family
    if track_id:
        e["revenant:track_id"] = int(track_id)
    if a.decoder:
        e["revenant:decoder"] = a.decoder
        e["revenant:decoded"] = {k: v for k, v in a.decoded}
        e["revenant:messages"] = int(a.messages)
    return e


def _stored(e: dict, where: str) -> StoredAnnotation:
    if not isinstance(e, dict):
        raise CaptureError(f"{where} is not an object")
    if "core:sample_start" not in e:
        raise CaptureError(f"{where} has no core:sample_start: it cannot be placed in the segment")

    def num(key, default, kind):
        v = e.get(key, default)
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise CaptureError(f"{where}: {key} is not a number")
        if kind is int and (v != int(v) or v < 0):
            raise CaptureError(f"{where}: {key} is not a count")
        return kind(v)

    def text(key):
        v = e.get(key, "")
        if not isinstance(v, str):
            raise CaptureError(f"{where}: {key} is not a string")
        return v

    decoded = e.get("revenant:decoded", {})
    if not isinstance(decoded, dict) or not all(isinstance(v, str) for v in decoded.values()):
        raise CaptureError(f"{where}: revenant:decoded is not an object of strings")
    a = Annotation(num("core:sample_start", 0, int), num("core:sample_count", 0, int),
                   num("core:freq_lower_edge", 0, int), num("core:freq_upper_edge", 0, int),
                   num("revenant:confidence", 0.0, float), num("revenant:peak_db", 0.0, float),
                   num("revenant:snr_db", 0.0, float), text("revenant:family"),
                   text("revenant:decoder"), list(decoded.items()),
                   num("revenant:messages", 0, int))
    return StoredAnnotation(a, text("core:generator"), num("revenant:track_id", 0, int))


def _load(text: str, where: str) -> dict:
    try:
        doc = json.loads(text)
    except json.JSONDecodeError as exc:
        raise CaptureError(f"{where} is not JSON: {exc}") from None
    if not isinstance(doc, dict):
        raise CaptureError(f"{where} is not a JSON object")
    return doc


def annotations_of(doc: dict, where: str = "the meta") -> list[StoredAnnotation]:
    arr = doc.get("annotations", [])
    if not isinstance(arr, list):
        raise CaptureError(f"{where}: annotations is not an array")
    return [_stored(e, f"{where} annotation {i}") for i, e in enumerate(arr)]


def stamps_of(doc: dict, where: str = "the meta") -> list[IndexStamp]:
    glob = doc.get("global", {})
    arr = glob.get(INDEXED_KEY, []) if isinstance(glob, dict) else []
    if not isinstance(arr, list):
        raise CaptureError(f"{where}: {INDEXED_KEY} is not an array")
    out = []
    for i, s in enumerate(arr):
        if not isinstance(s, dict) or not isinstance(s.get("generator"), str):
            raise CaptureError(f"{where}: {INDEXED_KEY} entry {i} names no generator")
        out.append(IndexStamp(s["generator"], str(s.get("at", "")), str(s.get("settings", ""))))
    return out


def read_annotations(meta_path: str) -> list[StoredAnnotation]:
    return annotations_of(_load(Path(meta_path).read_text(), meta_path), meta_path)


def read_index_stamps(meta_path: str) -> list[IndexStamp]:
    return stamps_of(_load(Path(meta_path).read_text(), meta_path), meta_path)


def indexed_by(stamps, annotations, generator: str = INDEX_GENERATOR) -> bool:
    """A pass under `generator` is on record: its stamp, or (older metas) an entry it left."""
    return any(s.generator == generator for s in stamps) or \
        any(a.generator == generator for a in annotations)


def _sorted(arr: list) -> list:
    return sorted(arr, key=lambda e: e.get("core:sample_start", 0))  # stable: ties keep order


def merge_annotations(meta_text: str, annotations: list[Annotation], generator: str,
                      stamp: IndexStamp | None = None) -> str:
    """The meta with `generator`'s entries replaced by `annotations`, ordered by
    sample_start, the other entries and keys kept; with a stamp, the pass is
    recorded in global revenant:indexed (replacing that generator's earlier one)."""
    doc = _load(meta_text, "the meta")
    annotations_of(doc)  # refuse a malformed array before touching it
    kept = [e for e in doc.get("annotations", []) if e.get("core:generator", "") != generator]
    doc["annotations"] = _sorted(kept + [_entry(a, generator) for a in annotations])
    if stamp is not None:
        glob = doc.get("global")
        if not isinstance(glob, dict):
            raise CaptureError("the meta has no global object to record the pass in")
        stamps_of(doc)
        others = [s for s in glob.get(INDEXED_KEY, []) if s.get("generator") != stamp.generator]
        glob[INDEXED_KEY] = others + [{"generator": stamp.generator, "at": stamp.at,
                                       "settings": stamp.settings}]
    return json.dumps(doc, indent=4)


def append_annotations(meta_text: str, annotations: list[StoredAnnotation]) -> str:
    """Adds entries, each under its own generator and track id, removing none."""
    doc = _load(meta_text, "the meta")
    annotations_of(doc)
    doc["annotations"] = _sorted(list(doc.get("annotations", [])) +
                                 [_entry(s.annotation, s.generator, s.track_id)
                                  for s in annotations])
    return json.dumps(doc, indent=4)


def _replace(meta_path: str, text: str) -> None:
    """A temporary beside the meta, renamed over it: a reader sees old or new."""
    d = os.path.dirname(os.path.abspath(meta_path))
    fd, tmp = tempfile.mkstemp(prefix=".meta-", dir=d)
    try:
        with os.fdopen(fd, "w") as f:
            f.write(text)
        os.replace(tmp, meta_path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def utc_now_iso() -> str:
    ns = time.time_ns()
    return time.strftime("%Y-%m-%dT%H:%M:%S", time.gmtime(ns // 10**9)) + \
        f".{(ns // 10**6) % 1000:03d}Z"


def write_annotations(meta_path: str, annotations: list[Annotation],
                      generator: str = INDEX_GENERATOR, settings: str = "") -> None:
    text = Path(meta_path).read_text()
    _replace(meta_path, merge_annotations(text, annotations, generator,
                                          IndexStamp(generator, utc_now_iso(), settings)))


def append_annotations_to(meta_path: str, annotations: list[StoredAnnotation]) -> None:
    _replace(meta_path, append_annotations(Path(meta_path).read_text(), annotations))


def indexed_segments(root: str | Path) -> list[IndexedSegment]:
    """Every closed segment under a capture root with its annotations and stamps."""
    out = []
    for e in catalog(root):
        doc = _load(Path(e.meta_path).read_text(), e.meta_path)
        out.append(IndexedSegment(e.data_path, e.info, annotations_of(doc, e.meta_path),
                                  stamps_of(doc, e.meta_path)))
    return out


# ---------------------------------------------------------------------------
# Queries
# ---------------------------------------------------------------------------

@dataclass
class Query:
    """Every field filters; unset passes. Time and frequency overlap, not contain."""
    from_ns: int | None = None
    to_ns: int | None = None
    freq_low: int | None = None
    freq_high: int | None = None
    min_confidence: float = 0.0
    min_snr_db: float | None = None
    min_duration_ns: int = 0
    family: str = ""       # case-insensitive substring; set excludes unlabelled entries
    generator: str = ""    # exact
    decoder: str = ""      # case-insensitive exact; set excludes non-decode entries
    decoded: list[tuple[str, str]] = field(default_factory=list)  # all must match


@dataclass
class Hit:
    start_ns: int
    end_ns: int
    freq_lower: int
    freq_upper: int
    data_path: str
    sample_start: int
    sample_count: int
    stream_epoch: int
    first_index: int
    confidence: float
    peak_db: float
    snr_db: float
    family: str
    generator: str
    track_id: int
    decoder: str
    decoded: list[tuple[str, str]]
    messages: int


def _span_ns(info: SegmentInfo, a: Annotation) -> tuple[int, int]:
    """Wall-clock start and end of an annotation (its first sample and one past
    its last, from the segment's anchor and rate)."""
    start = info.start_ns()
    if info.rate <= 0:
        return start, start
    first = start + a.sample_start * 1_000_000_000 // info.rate
    return first, start + (a.sample_start + a.sample_count) * 1_000_000_000 // info.rate


def _passes(s: StoredAnnotation, span: tuple[int, int], q: Query, time_freq: bool) -> bool:
    a = s.annotation
    if time_freq:
        if q.from_ns is not None and span[1] <= q.from_ns:
            return False
        if q.to_ns is not None and span[0] >= q.to_ns:
            return False
        if q.freq_low is not None and a.freq_upper < q.freq_low:
            return False
        if q.freq_high is not None and a.freq_lower > q.freq_high:
            return False
    if a.confidence < q.min_confidence:
        return False
    if q.min_snr_db is not None and a.snr_db < q.min_snr_db:
        return False
    if span[1] - span[0] < q.min_duration_ns:
        return False
    if q.family and q.family.lower() not in a.family.lower():
        return False
    if q.generator and s.generator != q.generator:
        return False
    if q.decoder and a.decoder.lower() != q.decoder.lower():
        return False
    if q.decoded:
        have = {(k.lower(), v.lower()) for k, v in a.decoded}
        if not all((k.lower(), v.lower()) in have for k, v in q.decoded):
            return False
    return True


def run_query(segments: list[IndexedSegment], q: Query) -> list[Hit]:
    """Every annotation passing the query, by start, then lower edge, then path."""
    hits = []
    for seg in segments:
        for s in seg.annotations:
            span = _span_ns(seg.info, s.annotation)
            if not _passes(s, span, q, True):
                continue
            a = s.annotation
            hits.append(Hit(span[0], span[1], a.freq_lower, a.freq_upper, seg.data_path,
                            a.sample_start, a.sample_count, seg.info.stream_epoch,
                            seg.info.first_index, a.confidence, a.peak_db, a.snr_db, a.family,
                            s.generator, s.track_id, a.decoder, list(a.decoded), a.messages))
    hits.sort(key=lambda h: (h.start_ns, h.freq_lower, h.data_path))
    return hits


@dataclass
class OccupancyRequest:
    from_ns: int
    to_ns: int
    bucket_ns: int
    freq_low: int
    freq_high: int
    filter: Query = field(default_factory=Query)


@dataclass
class OccupancyBucket:
    start_ns: int
    end_ns: int
    occupied: float   # share of the bucket with at least one annotation present
    captured: float   # share of the bucket any segment covered
    hits: int         # annotations overlapping the bucket


def _union_len(spans: list[tuple[int, int]], lo: int, hi: int) -> int:
    """Length of the union of spans clipped to [lo, hi)."""
    clipped = sorted((max(a, lo), min(b, hi)) for a, b in spans if b > lo and a < hi)
    total, end = 0, lo
    for a, b in clipped:
        a = max(a, end)
        if b > a:
            total += b - a
            end = b
    return total


def occupancy(segments: list[IndexedSegment], r: OccupancyRequest) -> list[OccupancyBucket]:
    if r.to_ns <= r.from_ns:
        raise CaptureError(f"an empty or inverted range [{r.from_ns}, {r.to_ns})")
    if r.bucket_ns <= 0:
        raise CaptureError(f"a bucket of {r.bucket_ns} ns")
    present, recorded = [], []
    for seg in segments:
        s0 = seg.info.start_ns()
        s1 = s0 + (seg.info.sample_count * 1_000_000_000 // seg.info.rate if seg.info.rate else 0)
        recorded.append((s0, s1))
        for s in seg.annotations:
            a = s.annotation
            if a.freq_upper < r.freq_low or a.freq_lower > r.freq_high:
                continue
            span = _span_ns(seg.info, a)
            if _passes(s, span, r.filter, False):
                present.append(span)
    out = []
    t = r.from_ns
    while t < r.to_ns:
        e = min(t + r.bucket_ns, r.to_ns)  # the last bucket stops at to_ns
        width = e - t
        out.append(OccupancyBucket(t, e, _union_len(present, t, e) / width,
                                   _union_len(recorded, t, e) / width,
                                   sum(1 for a, b in present if b > t and a < e)))
        t = e
    return out
