"""Detection events (rvx_candidates, rvx_tracker_*): the plain decision's
accepted bands as candidates in hertz, residuals of stopped signals withheld,
and tracks whose first and last sample indices bound each event
(core/detect/detector.cpp:1168-1900)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def event_config(bin_width_hz: float, bin_zero_hz: float, source_rate: int,
                 source_center_hz: int = 0, **kw) -> _lib.EventConfig:
    """DetectorConfig's defaults for the event stages, with the frame geometry."""
    c = _lib.EventConfig()
    _lib.lib().rvx_event_config_defaults(C.byref(c))
    c.bin_width_hz, c.bin_zero_hz = float(bin_width_hz), float(bin_zero_hz)
    c.source_rate, c.source_center_hz = int(source_rate), int(source_center_hz)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def candidates(cfg: _lib.EventConfig, average, floor, bands) -> list[_lib.Candidate]:
    """bands: [(start, width, deflection)] ascending and disjoint."""
    avg, flo = _f64(average), _f64(floor)
    arr = (_lib.DetectBand * max(1, len(bands)))(*[_lib.DetectBand(int(s), int(w), float(d))
                                                  for s, w, d in bands])
    cap = max(4, 4 * len(bands) + 4)
    out = (_lib.Candidate * cap)()
    n = C.c_uint32()
    _lib.check(_lib.lib().rvx_candidates(C.byref(cfg), avg.ctypes.data, flo.ctypes.data, avg.size,
                                         arr, len(bands), out, cap, C.byref(n)))
    return [out[i] for i in range(n.value)]


def measure_hd_sidebands(power, floor, bin_zero_hz: float, bin_width_hz: float,
                         centre_hz: float) -> _lib.HdReading:
    """The hybrid digital sideband reading around a wideband FM signal at
    centre_hz (core/detect/hd_hybrid.h), over a decision's average and floor."""
    p, f = _f64(power), _f64(floor)
    if p.size != f.size:
        raise ValueError("power and floor differ in length")
    out = _lib.HdReading()
    _lib.check(_lib.lib().rvx_measure_hd_sidebands(p.ctypes.data, f.ctypes.data, p.size,
                                                   float(bin_zero_hz), float(bin_width_hz),
                                                   float(centre_hz), C.byref(out)))
    return out


class Tracker:
    def __init__(self, cfg: _lib.EventConfig, bins: int):
        self.lib = _lib.lib()
        self.bins = bins
        h = C.c_void_p()
        _lib.check(self.lib.rvx_tracker_create(C.byref(cfg), bins, C.byref(h)))
        self.handle = h

    def update(self, average, floor, cands: list, now: int, elapsed_seconds: float) -> int:
        avg, flo = _f64(average), _f64(floor)
        arr = (_lib.Candidate * max(1, len(cands)))(*cands)
        kept = C.c_uint32()
        _lib.check(self.lib.rvx_tracker_update(self.handle, avg.ctypes.data, flo.ctypes.data,
                                               arr, len(cands), int(now), float(elapsed_seconds),
                                               C.byref(kept)))
        return kept.value

    def measure_levels(self, power_db) -> None:
        """One spectrum frame (dB per bin) into every live track's level
        statistics (Detector::measure_levels)."""
        p = np.ascontiguousarray(power_db, dtype=np.float32)
        _lib.check(self.lib.rvx_tracker_measure_levels(self.handle, p.ctypes.data_as(_lib.fptr),
                                                       p.size))

    def tracks(self) -> list[_lib.Track]:
        n = C.c_uint32()
        cap = 1024
        while True:
            out = (_lib.Track * cap)()
            rc = self.lib.rvx_tracker_tracks(self.handle, out, cap, C.byref(n))
            if rc == 0:
                return [out[i] for i in range(n.value)]
            if n.value <= cap:
                _lib.check(rc)
            cap = n.value

    def record_probe(self, track_id: int, family: int, double_sideband: bool = False) -> None:
        """A probe's family for a track (Detector::record_probe): FAMILY_UNKNOWN,
        FAMILY_UNMODULATED (with double_sideband for a DSB carrier) or FAMILY_OTHER."""
        _lib.check(self.lib.rvx_tracker_record_probe(self.handle, int(track_id), int(family),
                                                     int(bool(double_sideband))))

    def withdrawn_spurs(self) -> list[int]:
        """Absolute centres (Hz) of the spurs the latest decision took back."""
        n = C.c_uint32()
        cap = 64
        while True:
            out = (C.c_int64 * cap)()
            rc = self.lib.rvx_tracker_withdrawn_spurs(self.handle, out, cap, C.byref(n))
            if rc == 0:
                return [int(out[i]) for i in range(n.value)]
            if n.value <= cap:
                _lib.check(rc)
            cap = n.value

    def spur_book(self) -> list[_lib.SpurSighting]:
        """The sightings, to carry across a retune (Detector::spur_book)."""
        n = C.c_uint32()
        out = (_lib.SpurSighting * _lib.SPUR_BOOK_CAPACITY)()
        _lib.check(self.lib.rvx_tracker_spur_book(self.handle, out, _lib.SPUR_BOOK_CAPACITY,
                                                  C.byref(n)))
        return [out[i] for i in range(n.value)]

    def adopt_spur_book(self, book: list) -> None:
        """Takes a book from the tracker before a retune (Detector::adopt_spur_book)."""
        arr = (_lib.SpurSighting * max(1, len(book)))(*book)
        _lib.check(self.lib.rvx_tracker_adopt_spur_book(self.handle, arr, len(book)))

    def stats(self) -> _lib.TrackerStats:
        s = _lib.TrackerStats()
        _lib.check(self.lib.rvx_tracker_stats_get(self.handle, C.byref(s)))
        return s

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_tracker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
