"""The comb canceller's host half through the C ABI (rvx_comb_detect,
rvx_comb_bank_*), and the loop that runs it beside the engine's comb stage.

core/dsp/comb_cancel.h splits the stage in three: detection on a
time-averaged spectrum (about once a second), a tracker per comb that plans
each block's lines and reads back that block's projections, and the two
device passes. `CombCanceller` is the part of core/engine/graph.cpp that ties
them together around rvx_engine_push: before a push it asks the bank for the
push's lines (phases at the push's first sample) and hands them to
rvx_engine_set_comb; after it, it gives the push's projections back to the
bank.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

LOCK_NAMES = {_lib.COMB_IDLE: "idle", _lib.COMB_ACQUIRING: "acquiring",
              _lib.COMB_LOCKED: "locked", _lib.COMB_LOST: "lost",
              _lib.COMB_RECOVERING: "recovering"}


def _estimates(items) -> tuple:
    items = list(items or [])
    arr = (_lib.CombEstimate * max(1, len(items)))(*items)
    return arr, len(items)


def detect_combs(power: np.ndarray, rate: int, active=None, halves: bool = True) -> list:
    """Every comb in a linear power spectrum (natural order, DC at bins / 2),
    strongest first; `active` (CombEstimates) are held already and left out.
    halves=False reports a comb whole (detect_comb)."""
    p = np.ascontiguousarray(power, dtype=np.float32)
    act, n_act = _estimates(active)
    out = (_lib.CombEstimate * _lib.COMB_MAX_COMBS)()
    n = _lib.u32()
    _lib.check(_lib.lib().rvx_comb_detect(p.ctypes.data_as(_lib.fptr), p.size, int(rate), act,
                                          n_act, 1 if halves else 0, out, _lib.COMB_MAX_COMBS,
                                          C.byref(n)))
    return [out[i] for i in range(n.value)]


def detect_comb(power: np.ndarray, rate: int):
    """The strongest comb, whole, or None."""
    found = detect_combs(power, rate, halves=False)
    return found[0] if found else None


def spur(offset_hz: float) -> _lib.CombEstimate:
    e = _lib.CombEstimate()
    e.offset_hz, e.spur = float(offset_hz), 1
    return e


def take_spurs(bank, tracks, withdrawn_hz, source_center_hz: int, listened=()) -> None:
    """The detector's spur rule handed to the bank, once per decision
    (core/rpc/server.cpp:5661-5681, core/engine/graph.cpp:1615-1655): spurs
    inside a band an open receiver listens to go first, then the ones the
    decision withdrew, then every track named a spur is offered as a one-line
    estimate at its baseband offset unless a listened band holds it.
    tracks: the tracker's tracks; withdrawn_hz: its withdrawn centres
    (absolute Hz); listened: (low_hz, high_hz) baseband pairs, edges included."""
    bank.drop_spurs_in(listened)
    bank.withdraw_spurs([float(c - source_center_hz) for c in withdrawn_hz])
    offered = []
    for tr in tracks:
        if not tr.spur:
            continue
        f = float(tr.center_hz - source_center_hz)
        if any(lo <= f <= hi for lo, hi in listened):
            continue
        e = spur(f)
        e.lines_found = 1
        offered.append(e)
    bank.spurs(offered)


class CombBank:
    """Up to eight comb trackers (and sixteen spur trackers) behind one plan."""

    def __init__(self, rate: int):
        self.lib = _lib.lib()
        h = C.c_void_p()
        _lib.check(self.lib.rvx_comb_bank_create(int(rate), C.byref(h)))
        self.handle = h
        self.rate = int(rate)
        self._lines = (_lib.CombLine * _lib.COMB_MAX_LINES)()

    def detected(self, estimates) -> None:
        arr, n = _estimates(estimates)
        _lib.check(self.lib.rvx_comb_bank_detected(self.handle, arr, n))

    def spurs(self, estimates) -> None:
        arr, n = _estimates(estimates)
        _lib.check(self.lib.rvx_comb_bank_spurs(self.handle, arr, n))

    def drop_spurs(self) -> None:
        _lib.check(self.lib.rvx_comb_bank_drop_spurs(self.handle))

    def withdraw_spurs(self, frequencies_hz) -> None:
        """Every spur detected within 45 Hz of one of these offsets goes."""
        f = (C.c_double * max(1, len(frequencies_hz)))(*[float(x) for x in frequencies_hz])
        _lib.check(self.lib.rvx_comb_bank_withdraw_spurs(self.handle, f, len(frequencies_hz)))

    def drop_spurs_in(self, bands) -> None:
        """bands: (low_hz, high_hz) pairs, edges included."""
        b = (_lib.SpurBand * max(1, len(bands)))(*[_lib.SpurBand(lo, hi) for lo, hi in bands])
        _lib.check(self.lib.rvx_comb_bank_drop_spurs_in(self.handle, b, len(bands)))

    def reset(self) -> None:
        _lib.check(self.lib.rvx_comb_bank_reset(self.handle))

    def plan(self, start: int, count: int):
        """(ctypes array of CombLine, number of lines, subtract) for the block."""
        n, subs, sub = _lib.u32(), _lib.u32(), _lib.i32()
        _lib.check(self.lib.rvx_comb_bank_plan(self.handle, int(start), int(count), self._lines,
                                               _lib.COMB_MAX_LINES, C.byref(n), C.byref(subs),
                                               C.byref(sub)))
        lines = (_lib.CombLine * max(1, n.value))()
        C.memmove(lines, self._lines, C.sizeof(_lib.CombLine) * n.value)
        return lines, n.value, bool(sub.value)

    def observe(self, start: int, count: int, projections: np.ndarray) -> None:
        """projections: complex64 [lines, sub_blocks], the lines in plan order."""
        p = np.ascontiguousarray(projections, dtype=np.complex64)
        f = p.view(np.float32).ravel()
        _lib.check(self.lib.rvx_comb_bank_observe(self.handle, int(start), int(count),
                                                  p.shape[1] if p.ndim == 2 else 1,
                                                  f.ctypes.data_as(_lib.fptr), f.size))

    def status(self) -> _lib.CombStatus:
        s = _lib.CombStatus()
        _lib.check(self.lib.rvx_comb_bank_status(self.handle, C.byref(s)))
        return s

    def active(self) -> list:
        out = (_lib.CombEstimate * _lib.COMB_MAX_COMBS)()
        n = _lib.u32()
        _lib.check(self.lib.rvx_comb_bank_active(self.handle, out, _lib.COMB_MAX_COMBS,
                                                 C.byref(n)))
        return [out[i] for i in range(n.value)]

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_comb_bank_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class CancellerStats:
    pushes: int = 0
    detections: int = 0
    detections_found: int = 0


class CombCanceller:
    """A CombBank driving an engine's comb stage push by push. `detect(power)`
    takes a time-averaged span spectrum (DC at bins / 2) whenever one is
    ready; `push(codes)` plans the push's lines, pushes, and feeds the push's
    projections back."""

    def __init__(self, engine, rate: int, start_index: int = 0):
        self.engine = engine
        self.bank = CombBank(rate)
        self.rate = int(rate)
        self.next_index = int(start_index)
        self.stats = CancellerStats()
        # the stage opens before the first push with no lines
        engine.set_comb([], phase_index=self.next_index, subtract=False)

    def detect(self, power: np.ndarray) -> list:
        found = detect_combs(power, self.rate, self.bank.active())
        self.stats.detections += 1
        self.stats.detections_found += bool(found)
        self.bank.detected(found)
        return found

    def push(self, codes, count: int) -> None:
        """Pushes `count` samples (codes as rvx_engine_push takes them)."""
        start = self.next_index
        lines, n, subtract = self.bank.plan(start, count)
        self.engine.set_comb(lines if n else [], phase_index=start, subtract=subtract and n > 0)
        self.engine.push(codes)
        if n:
            proj, first, samples = self.engine.comb_projections()
            if first == start and samples == count and proj.shape[0] == n:
                self.bank.observe(start, count, proj)
        self.next_index += count
        self.stats.pushes += 1

    def status(self) -> _lib.CombStatus:
        return self.bank.status()
