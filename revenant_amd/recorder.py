"""The rolling capture behind a queue: the source's delivery thread hands
blocks over and returns at once, and one thread of the recorder's own does
every open, write, rename and delete (core/capture/writer.h, SegmentWriter).

Why a queue. A live source has to be drained on the device's schedule; a
write that stalls on a busy disk would become an overrun in the device. So
`consume` only copies the block into a buffer and queues it. The queue holds
`queue_blocks` blocks; a block that finds it full is dropped and counted,
and the samples it carried become a gap before the next block that gets in:
that block's segment starts after the gap with the gap in
revenant:dropped_before, and the segment before it closes as "loss". A
segment stays gapless whatever the disk does.

Who owns what. A queued block belongs to the recorder thread from the moment
it is queued. Counters and the queue sit under one lock; the file work runs
outside it, on the recorder thread only. An error stops the writing (it is
kept in `last_error`); the stream the recorder is attached to carries on,
and blocks offered after that are counted as unwritten.

Annotations arrive late: a detector reports a track when it ends, so a track
longer than what was left of its first segment arrives after that segment
closed. `add_annotation` takes the track's absolute stream span; the recorder
clips it to every segment of its stream epoch the span overlaps, each part
keeping the track id. The part in the open segment waits for that segment to
close; parts in segments already closed are appended to their metas through
a temporary and a rename. A part whose segment is gone (deleted by retention
or by hand), or older than the index of the last 1024 closed segments, is
counted as discarded.
"""
from __future__ import annotations

import collections
import threading
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

from .capture import CaptureError, CaptureWriter, SegmentInfo, SourceBlock
from .search import DETECT_GENERATOR, Annotation, StoredAnnotation, _entry, append_annotations_to

CLOSED_INDEX_LIMIT = 1024  # closed segments remembered for late annotation parts


@dataclass
class RecorderStats:
    """writer.h WriterStats."""
    segments_written: int = 0
    bytes_written: int = 0
    blocks_dropped: int = 0        # refused by a full queue
    samples_dropped: int = 0
    blocks_unwritten: int = 0      # offered after writing stopped on an error
    segments_deleted: int = 0
    annotations_discarded: int = 0
    annotations_merged_late: int = 0
    retained_bytes: int = 0
    last_error: str = ""


@dataclass
class _Queued:
    block: SourceBlock
    rate: int
    fmt: str
    epoch_anchor_ns: int


@dataclass
class _Part:
    """An annotation's span in absolute stream indices, waiting to be filed."""
    start: int
    count: int
    annotation: Annotation
    track_id: int


@dataclass
class _Closed:
    meta_path: Path
    first: int
    end: int


@dataclass
class SegmentRecorder:
    """A CaptureWriter run on its own thread behind a bounded queue.

    The writer's options (size and duration limits, retention by bytes and
    by age, the clock, the stream epoch, the hardware string, recovery at
    start) pass through unchanged."""
    root: str
    queue_blocks: int = 256
    max_segment_bytes: int = 256 << 20
    max_segment_seconds: float = 0.0
    retention_bytes: int = 0
    retention_max_age_ns: int = 0
    stream_epoch: int = 0
    hw: str = ""
    recover_on_start: bool = True
    now_ns: object = None
    _stats: RecorderStats = field(default_factory=RecorderStats, init=False)

    def __post_init__(self):
        if self.queue_blocks < 1:
            raise CaptureError(f"a queue of {self.queue_blocks} blocks holds nothing")
        self._lock = threading.Lock()
        self._wake = threading.Condition(self._lock)
        self._queue: collections.deque[_Queued] = collections.deque()
        self._incoming: list[_Part] = []
        self._pending_loss = 0      # samples lost since the last block queued
        self._gain: float | None = None
        self._running = False
        self._stopping = False
        self._errored = False
        self._thread: threading.Thread | None = None
        self._writer: CaptureWriter | None = None
        # recorder thread only
        self._waiting: list[_Part] = []
        self._closed: collections.deque[_Closed] = collections.deque()

    # -- the delivery side ---------------------------------------------------
    def start(self) -> None:
        """Recovers what a dead writer left open and starts the thread."""
        if self._running:
            return
        w = CaptureWriter(self.root, self.max_segment_bytes, self.retention_bytes,
                          self.stream_epoch, self.hw, self.recover_on_start,
                          self.max_segment_seconds, self.retention_max_age_ns, self.now_ns)
        w.annotations_for = self._annotations_for
        w.on_closed = self._on_closed
        w.gain_db = self._gain
        self.stream_epoch = w.stream_epoch
        self._writer = w
        self._stopping = False
        self._running = True
        self._thread = threading.Thread(target=self._run, name="capture-recorder", daemon=True)
        self._thread.start()

    def consume(self, block: SourceBlock, rate: int, fmt: str, epoch_anchor_ns: int = 0) -> None:
        """Never waits on the disk and never raises for it: a full queue or a
        stopped writer is counted instead."""
        with self._lock:
            if not self._running or self._stopping or self._errored:
                self._stats.blocks_unwritten += 1  # not started, stopping, or stopped on an error
                return
            if len(self._queue) >= self.queue_blocks:
                self._stats.blocks_dropped += 1
                self._stats.samples_dropped += block.samples
                # the samples it carried and any the device lost before it
                self._pending_loss += block.samples + block.dropped_before
                return
            lost, self._pending_loss = self._pending_loss, 0
        # the copy is made outside the lock: the caller's buffer is free on return
        copy = replace(block, codes=np.array(block.codes, copy=True),
                       dropped_before=block.dropped_before + lost)
        with self._lock:
            self._queue.append(_Queued(copy, rate, fmt, epoch_anchor_ns))
            self._wake.notify()

    def set_gain_db(self, gain_db: float | None) -> None:
        """The gain written as revenant:gain_db on segments that open after
        this call (a gain change does not close a segment)."""
        with self._lock:
            self._gain = gain_db
            if self._writer is not None:
                self._writer.gain_db = gain_db

    def add_annotation(self, absolute_start: int, annotation: Annotation,
                       track_id: int = 0) -> None:
        """`absolute_start` is the stream index of the first sample;
        annotation.sample_start is ignored. Safe from any thread."""
        with self._lock:
            self._incoming.append(_Part(int(absolute_start), int(annotation.sample_count),
                                        annotation, int(track_id)))
            self._wake.notify()

    def stop(self) -> None:
        """Drains the queue, closes the open segment with reason "stop",
        applies retention and joins. Safe to call twice."""
        with self._lock:
            if not self._running:
                return
            self._stopping = True
            self._wake.notify()
        self._thread.join()
        with self._lock:
            self._running = False

    def stats(self) -> RecorderStats:
        with self._lock:
            return replace(self._stats)

    # -- the recorder thread -------------------------------------------------
    def _run(self) -> None:
        while True:
            with self._lock:
                while not self._queue and not self._incoming and not self._stopping:
                    self._wake.wait()
                item = self._queue.popleft() if self._queue else None
                parts, self._incoming = self._incoming, []
                done = item is None and not parts and self._stopping
            if parts:
                self._route(parts)
            if item is not None:
                self._write(item)
            if done:
                break
        self._finish()

    def _write(self, item: _Queued) -> None:
        if self._errored:
            with self._lock:
                self._stats.blocks_unwritten += 1
            return
        w = self._writer
        try:
            w.write(item.block, item.rate, item.fmt, item.epoch_anchor_ns)
        except (OSError, CaptureError) as exc:
            self._fail(f"writing the block at index {item.block.first_index}: {exc}")
            return
        self._publish()

    def _finish(self) -> None:
        if not self._errored:
            try:
                self._writer.close("stop")
            except (OSError, CaptureError) as exc:
                self._fail(f"closing the last segment: {exc}")
        # parts still waiting lie past every segment that was written
        with self._lock:
            self._stats.annotations_discarded += len(self._waiting)
        self._waiting = []
        self._publish()

    def _fail(self, message: str) -> None:
        with self._lock:
            self._errored = True
            self._stats.last_error = message
            self._stats.blocks_unwritten += 1

    def _publish(self) -> None:
        w = self._writer
        with self._lock:
            s = self._stats
            s.segments_written = w.segments_written
            s.bytes_written = w.bytes_written
            s.segments_deleted = w.segments_deleted
            s.retained_bytes = w.retained_bytes

    # -- annotations (recorder thread) -------------------------------------------
    def _route(self, parts: list[_Part]) -> None:
        """Files each part's overlap with every closed segment now; keeps the
        rest for the segment that will hold it."""
        merged = discarded = 0
        floor = self._closed[0].first if self._closed else None
        for p in parts:
            lo, hi = p.start, p.start + p.count
            if floor is not None and lo < floor:
                discarded += 1  # older than the closed-segment index reaches
                lo = floor
            for c in self._closed:
                a, b = max(lo, c.first), min(hi, c.end)
                if a >= b and not (p.count == 0 and c.first <= lo < c.end):
                    continue
                piece = self._piece(p, a - c.first, max(b - a, 0))
                try:
                    append_annotations_to(str(c.meta_path), [piece])
                    merged += 1
                except (OSError, CaptureError):
                    discarded += 1  # the segment is gone, or its meta cannot be replaced
            end = self._closed[-1].end if self._closed else None
            if end is None or hi > end or (p.count == 0 and lo >= end):
                self._waiting.append(_Part(max(lo, end or lo), hi - max(lo, end or lo),
                                           p.annotation, p.track_id))
        with self._lock:
            self._stats.annotations_merged_late += merged
            self._stats.annotations_discarded += discarded

    @staticmethod
    def _piece(p: _Part, start: int, count: int) -> StoredAnnotation:
        a = replace(p.annotation, sample_start=start, sample_count=count)
        return StoredAnnotation(a, DETECT_GENERATOR, p.track_id)

    def _annotations_for(self, info: SegmentInfo) -> list[dict]:
        """The entries of a segment closing now: every waiting part's overlap
        with it; parts reaching past it keep waiting for the next."""
        lo_seg, hi_seg = info.first_index, info.end_index()
        out, keep = [], []
        for p in self._waiting:
            a, b = max(p.start, lo_seg), min(p.start + p.count, hi_seg)
            if a < b or (p.count == 0 and lo_seg <= p.start < hi_seg):
                piece = self._piece(p, a - lo_seg, max(b - a, 0))
                out.append(_entry(piece.annotation, piece.generator, piece.track_id))
            if p.start + p.count > hi_seg:
                keep.append(p)
            elif p.start + p.count <= lo_seg and p.count > 0:
                with self._lock:  # it lay in a gap or before this epoch's segments
                    self._stats.annotations_discarded += 1
        self._waiting = keep
        return out

    def _on_closed(self, meta_path: Path, info: SegmentInfo) -> None:
        if self._closed and self._closed[-1].end > info.first_index:
            self._closed.clear()  # a new stream epoch: indices restart
        self._closed.append(_Closed(meta_path, info.first_index, info.end_index()))
        while len(self._closed) > CLOSED_INDEX_LIMIT:
            self._closed.popleft()
