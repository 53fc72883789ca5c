"""Front-end correction on the host: the estimator that turns a block's chunk
sums into the DC and I/Q-imbalance estimate and the correction's numbers
(core/dsp/front_end_correction.h FrontEndCorrector, FrontEndEstimate,
total_moments). The arithmetic is the library's (front_end.cpp); the per-sample
work is on the device (Device.iq_moments, Device.iq_correct, and the chain's
fused correction).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

CHUNK = 64          # RVX_IQ_MOMENTS_CHUNK
PER_CHUNK = 5       # RVX_IQ_MOMENTS_PER_CHUNK


def moments_chunks(count: int) -> int:
    """front_end_correction.h:90 iq_moments_chunks."""
    return (count + CHUNK - 1) // CHUNK


def total_moments(sums: np.ndarray, count: int) -> _lib.IqMomentsTotal:
    """front_end_correction.cpp:39 total_moments (chunk order, in double)."""
    s = np.ascontiguousarray(sums, dtype=np.float32).ravel()
    out = _lib.IqMomentsTotal()
    _lib.check(_lib.lib().rvx_total_moments(s.ctypes.data_as(_lib.fptr), s.size, count,
                                            C.byref(out)))
    return out


@dataclass
class Estimate:
    """FrontEndEstimate with its derived readings."""

    measured: bool
    iq_plausible: bool
    samples_seen: int
    dc_i: float
    dc_q: float
    gain: float
    sin_phase: float
    dc_dbfs: float
    gain_error_db: float
    phase_error_deg: float
    image_rejection_db: float


class FrontEndCorrector:
    """core/dsp/front_end_correction.h:222-247."""

    def __init__(self, dc_time_constant_s: float = 0.1, iq_time_constant_s: float = 1.0,
                 max_gain_error_db: float = 6.0, max_phase_error_deg: float = 30.0):
        self.lib = _lib.lib()
        cfg = _lib.FrontEndConfig(dc_time_constant_s, iq_time_constant_s, max_gain_error_db,
                                  max_phase_error_deg)
        h = C.c_void_p()
        _lib.check(self.lib.rvx_front_end_create(C.byref(cfg), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_front_end_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> None:
        self.lib.rvx_front_end_reset(self.handle)

    def update(self, moments: _lib.IqMomentsTotal, rate: int) -> None:
        _lib.check(self.lib.rvx_front_end_update(self.handle, C.byref(moments), rate))

    def update_sums(self, sums: np.ndarray, count: int, rate: int) -> None:
        self.update(total_moments(sums, count), rate)

    def estimate(self) -> Estimate:
        e = _lib.FrontEndEstimate()
        _lib.check(self.lib.rvx_front_end_estimate_get(self.handle, C.byref(e)))
        d = [C.c_double() for _ in range(4)]
        _lib.check(self.lib.rvx_front_end_derived(C.byref(e), *[C.byref(x) for x in d]))
        return Estimate(bool(e.measured), bool(e.iq_plausible), e.samples_seen, e.dc_i, e.dc_q,
                        e.gain, e.sin_phase, *[x.value for x in d])

    def correction(self, dc: bool = True, iq: bool = True) -> _lib.IqCorrectParams:
        """The kernel's numbers (ring fields zero); the identity for a half
        that is off, not yet measured, or implausible."""
        out = _lib.IqCorrectParams()
        _lib.check(self.lib.rvx_front_end_correction(self.handle, int(dc), int(iq),
                                                     C.byref(out)))
        return out
