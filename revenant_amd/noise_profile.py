"""The adaptive noise profile through the C ABI (rvx_noise_profile_*).

core/dsp/noise_profile.h's NoiseProfile: one receiver's steady noise (hum,
whine, birdies) learned and taken out of its mono audio on the completion
thread, after the audio is read back (core/engine/graph.cpp:2587-2592 runs one
per receiver slot, with `quiet` = the squelch is shut). Host arithmetic only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

MAX_STRENGTH_DB = _lib.NOISE_PROFILE_MAX_STRENGTH_DB


@dataclass
class NoiseProfileSettings:
    """NoiseProfileSettings (noise_profile.h:44-50)."""
    enabled: bool = False
    strength_db: float = 12.0
    auto_learn: bool = True
    learn_count: int = 0
    reset_count: int = 0

    def _c(self) -> _lib.NoiseProfileSettings:
        return _lib.NoiseProfileSettings(int(self.enabled), int(self.auto_learn),
                                         float(self.strength_db), int(self.learn_count),
                                         int(self.reset_count))


class NoiseProfile:
    def __init__(self):
        self.lib = _lib.lib()
        h = C.c_void_p()
        _lib.check(self.lib.rvx_noise_profile_create(C.byref(h)))
        self.handle = h

    def process(self, audio: np.ndarray, rate: float, quiet: bool,
                settings: NoiseProfileSettings) -> None:
        """Mono float32 audio in place (a contiguous array)."""
        if audio.dtype != np.float32 or not audio.flags.c_contiguous:
            raise ValueError("audio must be a contiguous float32 array")
        s = settings._c()
        _lib.check(self.lib.rvx_noise_profile_process(
            self.handle, audio.ctypes.data_as(_lib.fptr), audio.size, float(rate),
            1 if quiet else 0, C.byref(s)))

    def retune(self) -> None:
        _lib.check(self.lib.rvx_noise_profile_retune(self.handle))

    def _state(self):
        learning, learned, latency, bins = _lib.i32(), _lib.i32(), _lib.u32(), _lib.u32()
        _lib.check(self.lib.rvx_noise_profile_state(self.handle, C.byref(learning),
                                                    C.byref(learned), C.byref(latency),
                                                    C.byref(bins)))
        return bool(learning.value), bool(learned.value), latency.value, bins.value

    def learning(self) -> bool:
        return self._state()[0]

    def learned(self) -> bool:
        return self._state()[1]

    def latency(self) -> int:
        """Samples between a sample going in and its coming out, while enabled."""
        return self._state()[2]

    def profile(self) -> np.ndarray:
        """The estimate in use, power per bin; empty until there is one."""
        bins = self._state()[3]
        out = np.zeros(bins, dtype=np.float64)
        n = _lib.u32()
        _lib.check(self.lib.rvx_noise_profile_estimate(
            self.handle, out.ctypes.data_as(C.POINTER(C.c_double)), bins, C.byref(n)))
        return out[:n.value]

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_noise_profile_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
