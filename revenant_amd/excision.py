"""Impulse excision on the host side: the design, the ring sizes, the
validation and the cursor that plans each block's four passes
(core/dsp/impulse_excision.h ExcisionConfig, design_excision,
excision_threshold, size_excision_rings, validate, ExcisionCursor). The
arithmetic is the library's (excision.hip); the passes run on the device
through Device.excise_*.
"""
from __future__ import annotations

import ctypes as C

from . import _lib

SEGMENT, REFERENCE, FLAG, APPLY = 0, 1, 2, 3   # RVX_EXCISE_*
COUNTERS = 4                                  # samples, events, spared, unused
DEFAULT_THRESHOLD_DB = 15.0
DEFAULT_MAX_WIDTH_S = 20e-6


def design_excision(source_rate: int,
                    max_width_seconds: float = DEFAULT_MAX_WIDTH_S) -> _lib.ExcisionConfig:
    out = _lib.ExcisionConfig()
    _lib.check(_lib.lib().rvx_design_excision(source_rate, max_width_seconds, C.byref(out)))
    return out


def excision_threshold(threshold_db: float = DEFAULT_THRESHOLD_DB) -> float:
    return float(_lib.lib().rvx_excision_threshold(threshold_db))


def size_excision_rings(config: _lib.ExcisionConfig, max_block: int, frames_in_flight: int,
                        ring_mask: int) -> _lib.ExcisionRings:
    out = _lib.ExcisionRings()
    _lib.check(_lib.lib().rvx_size_excision_rings(C.byref(config), max_block, frames_in_flight,
                                                  ring_mask, C.byref(out)))
    return out


def check(config: _lib.ExcisionConfig, params: _lib.ExcisionParams, pass_: int) -> None:
    c = _lib.ExcisionConfig.from_buffer_copy(bytes(config))
    c.pass_ = pass_
    _lib.check(_lib.lib().rvx_excision_check(C.byref(c), C.byref(params)))


class ExcisionCursor:
    """Which samples each pass covers, block by block (impulse_excision.h:254-283)."""

    def __init__(self, config: _lib.ExcisionConfig, ring_mask: int, rings: _lib.ExcisionRings,
                 threshold: float):
        self.lib = _lib.lib()
        h = C.c_void_p()
        _lib.check(self.lib.rvx_excision_cursor_create(C.byref(config), ring_mask, C.byref(rings),
                                                       threshold, C.byref(h)))
        self.handle = h

    def plan(self, start: int, count: int) -> _lib.ExcisionPlan:
        out = _lib.ExcisionPlan()
        _lib.check(self.lib.rvx_excision_cursor_plan(self.handle, start, count, C.byref(out)))
        return out

    def reset(self) -> None:
        self.lib.rvx_excision_cursor_reset(self.handle)

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_excision_cursor_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
