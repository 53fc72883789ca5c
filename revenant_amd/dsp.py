"""Host-side design vocabulary, mirroring core/dsp's planners.

Names and argument meanings follow the reference (core/dsp/pfb.h,
core/dsp/spectrum_reference.h, core/dsp/vrx_reference.h). The arithmetic is
the library's (revenant_amd/csrc/design.cpp), so the GPU path and every test
read exactly the same float tables.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

FORMAT_CU8, FORMAT_CS8, FORMAT_CS16, FORMAT_CS24 = 0, 1, 2, 3
BYTES_PER_SAMPLE = {FORMAT_CU8: 2, FORMAT_CS8: 2, FORMAT_CS16: 4, FORMAT_CS24: 6}

DEMOD_RAW, DEMOD_AM, DEMOD_NFM, DEMOD_WFM, DEMOD_USB, DEMOD_LSB, DEMOD_DSB, DEMOD_CW = range(8)


@dataclass(frozen=True)
class GridParams:
    """core/dsp/pfb.h:32-74. taps_per_branch is L + 1 (the delay pad included)."""

    channels: int = 64
    taps_per_branch: int = 17
    decimation: int = 32

    @property
    def prototype_length(self) -> int:
        return self.channels * self.taps_per_branch

    @property
    def stages(self) -> int:
        return self.channels.bit_length() - 1

    def c(self) -> _lib.Grid:
        return _lib.Grid(self.channels, self.taps_per_branch, self.decimation)


@dataclass
class PrototypeFilter:
    """core/dsp/pfb.h PrototypeFilter."""

    taps: np.ndarray
    stopband_db: float
    group_delay_samples: int


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def design_prototype(grid: GridParams, attenuation_db: float = 120.0,
                     beta: float | None = None) -> PrototypeFilter:
    """core/dsp/pfb_design.cpp:381-510. ``beta`` overrides the Kaiser shape
    parameter that ``attenuation_db`` would select."""
    out = np.zeros(grid.prototype_length, dtype=np.float32)
    sb = C.c_double()
    gd = C.c_uint64()
    g = grid.c()
    _lib.check(_lib.lib().rvx_design_prototype(C.byref(g), attenuation_db,
                                               -1.0 if beta is None else float(beta),
                                               _fptr(out), out.size, C.byref(sb), C.byref(gd)))
    return PrototypeFilter(out, sb.value, gd.value)


def build_twiddles(channels: int) -> np.ndarray:
    """core/dsp/pfb_design.cpp:515-591; returns complex64 W_M^j."""
    out = np.zeros(2 * channels, dtype=np.float32)
    _lib.check(_lib.lib().rvx_build_twiddles(channels, _fptr(out), out.size))
    return out.view(np.complex64)


def build_spectrum_window(size: int, taps_per_branch: int = 17,
                          attenuation_db: float = 120.0) -> np.ndarray:
    """core/dsp/spectrum_reference.cpp:195-285: N window taps then N/2 gains."""
    out = np.zeros(size + size // 2, dtype=np.float32)
    _lib.check(_lib.lib().rvx_build_spectrum_window(size, taps_per_branch, attenuation_db,
                                                    _fptr(out), out.size))
    return out


def build_nco_table(log2_size: int = 16) -> np.ndarray:
    """core/dsp/vrx_reference.cpp:364; exp(-j 2 pi i / K), complex64."""
    out = np.zeros(2 << log2_size, dtype=np.float32)
    _lib.check(_lib.lib().rvx_build_nco_table(log2_size, _fptr(out), out.size))
    return out.view(np.complex64)


def turn_fixed64(numerator: int, denominator: int) -> int:
    """core/dsp/vrx_reference.cpp:308: round(numerator / denominator * 2^64)."""
    return int(_lib.lib().rvx_turn_fixed64(numerator, denominator))


def nco_delta(hz_numerator: int, hz_denominator: int, rate: int) -> int:
    """core/dsp/vrx_reference.cpp:333: the 0.64 phase step of f = num/den Hz."""
    out = C.c_uint64()
    _lib.check(_lib.lib().rvx_nco_delta(hz_numerator, hz_denominator, rate, C.byref(out)))
    return out.value


def ema_frame_decay(alpha: float, frame_blocks: int) -> np.float32:
    """(1 - alpha)^frame_blocks rounded once to float: the frame recurrence's factor."""
    return np.float32((1.0 - float(np.float32(alpha))) ** frame_blocks)
