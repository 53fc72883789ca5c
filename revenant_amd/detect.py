"""The plain detector's decision (rvx_detector / rvx_detect): noise floor,
summed-bin ladder and greedy growth of core/detect/detector.cpp:794-1120 over
an averaged linear spectrum on the device."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .ops import Device, _p


def config(bins: int, bin_width_hz: float, threshold_db: float = 6.0, **kw) -> _lib.DetectConfig:
    """DetectorConfig's defaults (core/detect/detector.h:579-909)."""
    return _lib.DetectConfig(bins, kw.get("noise_knots", 32), kw.get("noise_window_knots", 8),
                             kw.get("max_sum_bins", 0), kw.get("max_peaks", 0),
                             kw.get("split_gap_bins", 8), bin_width_hz, threshold_db,
                             kw.get("noise_percentile", 0.10),
                             kw.get("noise_percentile_high", 0.40),
                             kw.get("noise_excision_sigma", 4.0),
                             kw.get("edge_floor_fraction", 0.0), kw.get("edge_floor_sigma", 3.0))


class Detector:
    def __init__(self, dev: Device, cfg: _lib.DetectConfig):
        self.dev = dev
        self.cfg = cfg
        h = C.c_void_p()
        _lib.check(dev.lib.rvx_detector_create(dev.ctx, C.byref(cfg), C.byref(h)))
        self.handle = h
        self._bands = (_lib.DetectBand * cfg.bins)()

    def decide(self, average: torch.Tensor, allow_prefix: torch.Tensor | None = None):
        """average: CUDA float64 [bins]; allow_prefix: CUDA int32 [bins + 1] or None.
        Returns ([(start, width, deflection)], peaks found)."""
        if average.dtype != torch.float64 or average.numel() != self.cfg.bins:
            raise ValueError("average must be float64 with one value per bin")
        if allow_prefix is not None and (allow_prefix.dtype != torch.int32 or
                                         allow_prefix.numel() != self.cfg.bins + 1):
            raise ValueError("allow_prefix must be int32 with bins + 1 entries")
        nb, npk = _lib.u32(), _lib.u32()
        _lib.check(self.dev.lib.rvx_detect(self.handle, _p(average), _p(allow_prefix),
                                           self._bands, self.cfg.bins, C.byref(nb),
                                           C.byref(npk)))
        return [(b.start, b.width, b.deflection) for b in self._bands[:nb.value]], npk.value

    def floor(self):
        """The last decision's per-bin floor and spread (host float64 copies)."""
        import numpy as np

        out = []
        for fn in (self.dev.lib.rvx_detector_floor, self.dev.lib.rvx_detector_sigma):
            a = np.empty(self.cfg.bins, dtype=np.float64)
            _lib.check(self.dev.lib.rvx_buf_read(self.dev.ctx, a.ctypes.data, fn(self.handle),
                                                 a.nbytes))
            out.append(a)
        return out[0], out[1]

    def close(self) -> None:
        if self.handle:
            self.dev.lib.rvx_detector_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
