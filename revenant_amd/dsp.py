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
FORMAT_CF32 = 4  # converted samples: the chain's input after conditioning
BYTES_PER_SAMPLE = {FORMAT_CU8: 2, FORMAT_CS8: 2, FORMAT_CS16: 4, FORMAT_CS24: 6}

DEMOD_RAW, DEMOD_AM, DEMOD_NFM, DEMOD_WFM, DEMOD_USB, DEMOD_LSB, DEMOD_DSB, DEMOD_CW = range(8)
DEMOD_SAM = 12
# core/engine/vrx.h:213 Deemphasis: DEFAULT takes the mode's own curve.
DEEMPH_DEFAULT, DEEMPH_NONE, DEEMPH_US75, DEEMPH_EU50 = range(4)
CARRIER_NONE, CARRIER_PLL, CARRIER_COSTAS = 0, 1, 2  # core/dsp/vrx_reference.h:674-676
CARRIER_FLAG_RESET = 1
CARRIER_STATE_WORDS = 8
# core/dsp/vrx_reference.h:787-820, the loop's design figures
CARRIER_ACQUIRE_HZ, CARRIER_TRACK_HZ = 400.0, 30.0
CARRIER_DAMPING = 0.70710678118654752440


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


def ema_params(channels: int, chan_blocks: int, in_offset: int, frame_blocks: int, frames: int,
               alpha: float, row_frames: int = 1) -> "_lib.EmaParams":
    """The averaged-spectrum parameter block with both decay factors filled in."""
    a = np.float32(alpha)
    return _lib.EmaParams(channels, chan_blocks, chan_blocks - 1, in_offset, frame_blocks,
                          frames, a, ema_frame_decay(a, frame_blocks), row_frames,
                          ema_frame_decay(a, frame_blocks * row_frames))


# ---------------------------------------------------------------------------
# Receivers (core/engine/vrx_place.cpp, core/dsp/vrx_reference.cpp)
# ---------------------------------------------------------------------------

@dataclass
class VrxRequest:
    """core/engine/vrx.h VrxParams: what a receiver asks for."""

    center_hz: int
    mode: int
    bandwidth_hz: int = 0
    audio_rate: int = 48000
    cw_pitch: int = 0
    passband: tuple[int, int] | None = None
    deemphasis: int = DEEMPH_DEFAULT
    stereo: bool = True   # a request: kept by WFM programme audio only

    def c(self) -> _lib.VrxRequest:
        lo, hi = self.passband if self.passband else (0, 0)
        return _lib.VrxRequest(self.center_hz, self.bandwidth_hz, lo, hi, self.mode,
                               self.deemphasis, self.audio_rate, self.cw_pitch,
                               0 if self.stereo else 1, 0)


@dataclass
class VrxPlan:
    """A planned receiver: the C plan block and its float tables."""

    plan: _lib.VrxPlan
    fine_taps: np.ndarray  # complex64, phases*taps + 1
    weights: np.ndarray    # float32: audio_taps + dc_taps (+ stereo pilot and rotation pairs)

    @property
    def fine(self) -> _lib.VrxFineConfig:
        return self.plan.fine

    @property
    def demod(self) -> _lib.VrxDemodConfig:
        return self.plan.demod

    def fine_block(self, channel_base: int, channel_mask: int, fine_mask: int,
                   first_output: int, count: int) -> tuple[_lib.VrxFineParams, int]:
        out = _lib.VrxFineParams()
        first_input = C.c_uint64()
        _lib.check(_lib.lib().rvx_fine_block(C.byref(self.plan), channel_base, channel_mask,
                                             fine_mask, first_output, count, C.byref(out),
                                             C.byref(first_input)))
        return out, first_input.value

    def demod_block(self, fine_mask: int, first_audio: int, count: int) -> _lib.VrxDemodParams:
        out = _lib.VrxDemodParams()
        _lib.check(_lib.lib().rvx_demod_block(C.byref(self.plan), fine_mask, first_audio, count,
                                              C.byref(out)))
        return out

    @property
    def history(self) -> int:
        """Fine samples below an output's index the demodulator reads."""
        return int(_lib.lib().rvx_demod_history(C.byref(self.plan.demod)))

    @property
    def carrier(self) -> _lib.VrxCarrierConfig:
        return self.plan.carrier

    def carrier_block(self, fine_mask: int, first_fine: int, count: int, reset: bool = False,
                      shift: float = 0.0) -> _lib.VrxCarrierParams:
        """core/dsp/vrx_reference.cpp:2911 carrier_block."""
        out = _lib.VrxCarrierParams()
        _lib.check(_lib.lib().rvx_carrier_block(C.byref(self.plan), fine_mask, first_fine, count,
                                                1 if reset else 0, shift, C.byref(out)))
        return out


def carrier_power(mode: int) -> int:
    """core/dsp/vrx_reference.cpp:918 carrier_power."""
    return int(_lib.lib().rvx_carrier_power(mode))


def carrier_loop_gains(noise_bandwidth_hz: float, damping: float, rate: int) -> tuple[float, float]:
    """core/dsp/vrx_reference.cpp:948 carrier_loop_gains: (K1, K2) in double."""
    k1, k2 = C.c_double(), C.c_double()
    _lib.check(_lib.lib().rvx_carrier_loop_gains(noise_bandwidth_hz, damping, rate, C.byref(k1),
                                                 C.byref(k2)))
    return k1.value, k2.value


def design_carrier_loop(demod_rate: int, reach_hz: int) -> _lib.VrxCarrierParams:
    """core/dsp/vrx_reference.cpp:958 design_carrier_loop."""
    out = _lib.VrxCarrierParams()
    _lib.check(_lib.lib().rvx_design_carrier_loop(demod_rate, reach_hz, C.byref(out)))
    return out


def carrier_retune(moved_hz: int, granted_width: int, demod_rate: int) -> tuple[bool, float]:
    """core/dsp/vrx_reference.cpp:1187 carrier_retune: (reset, shift)."""
    reset, shift = C.c_int32(), C.c_double()
    _lib.check(_lib.lib().rvx_carrier_retune(moved_hz, granted_width, demod_rate, C.byref(reset),
                                             C.byref(shift)))
    return bool(reset.value), shift.value


def carrier_check(config: _lib.VrxCarrierConfig, params: _lib.VrxCarrierParams) -> None:
    """core/dsp/vrx_reference.cpp:986 validate; raises RvxError with its message."""
    _lib.check(_lib.lib().rvx_vrx_carrier_check(C.byref(config), C.byref(params)))


def channel_count_for(rate: int, widest_receiver_hz: int) -> int:
    """core/engine/vrx_place.cpp:187: the channel count whose channels carry a
    receiver of the given width wherever it lands."""
    return int(_lib.lib().rvx_channel_count_for(int(rate), int(widest_receiver_hz)))


def channel_centre(grid: GridParams, rate: int, channel: int) -> tuple[int, int]:
    """core/dsp/pfb_design.cpp:349: (numerator, denominator) Hz, reduced."""
    num, den = C.c_int64(), C.c_int64()
    g = grid.c()
    _lib.check(_lib.lib().rvx_channel_centre(C.byref(g), int(rate), int(channel), C.byref(num),
                                             C.byref(den)))
    return num.value, den.value


def channel_sample_to_input(grid: GridParams, prototype: PrototypeFilter, base: int,
                            m: int) -> int:
    """core/dsp/pfb_design.cpp:593: the input index channel sample m is centred
    on (modulo 2^64)."""
    g = grid.c()
    return int(_lib.lib().rvx_channel_sample_to_input(
        C.byref(g), int(prototype.group_delay_samples), int(base) % 2**64, int(m) % 2**64))


def resolve_passband(request: VrxRequest) -> tuple[int, int]:
    """core/dsp/vrx_reference.h resolve_passband."""
    lo, hi = C.c_int64(), C.c_int64()
    r = request.c()
    _lib.check(_lib.lib().rvx_resolve_passband(C.byref(r), C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def default_passband(mode: int) -> tuple[int, int]:
    """core/dsp/vrx_reference.h default_passband."""
    lo, hi = C.c_int64(), C.c_int64()
    _lib.check(_lib.lib().rvx_default_passband(mode, C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def max_channel_bandwidth(placement: _lib.VrxPlacement) -> int:
    """core/dsp/vrx_reference.cpp:1409."""
    return int(_lib.lib().rvx_max_channel_bandwidth(C.byref(placement)))


def clamp_to_channel(placement: _lib.VrxPlacement, band: tuple[int, int]) -> tuple[int, int]:
    """core/dsp/vrx_reference.cpp:1359."""
    lo, hi = C.c_int64(band[0]), C.c_int64(band[1])
    _lib.check(_lib.lib().rvx_clamp_to_channel(C.byref(placement), C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def minimum_demod_rate(mode: int, bandwidth: int, cw_pitch: int = 0) -> int:
    """core/dsp/vrx_reference.cpp:1519, the bandwidth shorthand overload."""
    return int(_lib.lib().rvx_minimum_demod_rate(mode, bandwidth, cw_pitch))


def demod_rate_for(request: VrxRequest, placement: _lib.VrxPlacement, audio_rate: int) -> int:
    """core/dsp/vrx_reference.cpp:1670: the rate plan_vrx builds, cheaply."""
    out = C.c_int64()
    r = request.c()
    _lib.check(_lib.lib().rvx_demod_rate_for(C.byref(r), C.byref(placement), audio_rate,
                                             C.byref(out)))
    return out.value


def place(grid: GridParams, rate: int, request: VrxRequest) -> _lib.VrxPlacement:
    """core/engine/vrx_place.cpp:216."""
    out = _lib.VrxPlacement()
    g, r = grid.c(), request.c()
    _lib.check(_lib.lib().rvx_place(C.byref(g), rate, C.byref(r), C.byref(out)))
    return out


def plan_vrx(grid: GridParams, rate: int, request: VrxRequest,
             placement: _lib.VrxPlacement | None = None) -> VrxPlan:
    """core/dsp/vrx_reference.cpp:2568 (the modes this build plans)."""
    if placement is None:
        placement = place(grid, rate, request)
    out = _lib.VrxPlan()
    g, r = grid.c(), request.c()
    _lib.check(_lib.lib().rvx_plan_vrx(C.byref(g), rate, C.byref(r), C.byref(placement),
                                       C.byref(out)))
    n = out.fine.phases * out.fine.taps + 1
    taps = np.zeros(2 * n, dtype=np.float32)
    d = out.demod
    stereo = 2 * d.pilot_taps + 2 * d.audio_taps if d.pilot_taps else 0
    w = np.zeros(d.audio_taps + d.dc_taps + stereo, dtype=np.float32)
    _lib.check(_lib.lib().rvx_plan_vrx_tables(C.byref(out), _fptr(taps), taps.size, _fptr(w),
                                              w.size))
    return VrxPlan(out, taps.view(np.complex64), w)


def design_deemphasis(curve: int, audio_rate: int) -> np.ndarray:
    """vrx_reference.cpp:2011 design_deemphasis_taps: the resolved curve's float taps."""
    n = C.c_uint32()
    _lib.check(_lib.lib().rvx_design_deemphasis(curve, audio_rate, None, 0, C.byref(n)))
    taps = np.zeros(n.value, np.float32)
    _lib.check(_lib.lib().rvx_design_deemphasis(curve, audio_rate, _fptr(taps), taps.size,
                                                C.byref(n)))
    return taps


def prescale_taps(taps: np.ndarray) -> np.ndarray:
    """taps * 2^-15 exactly, for a cs16 chain with taps_scaled = 1
    (rvx_prescale_taps refuses taps that would leave the normal range)."""
    t = np.ascontiguousarray(taps, dtype=np.float32)
    out = np.empty_like(t)
    _lib.check(_lib.lib().rvx_prescale_taps(_fptr(t), t.size, _fptr(out)))
    return out


# ---------------------------------------------------------------------------
# Comb cancellation (core/dsp/comb_cancel.h): the device half's vocabulary
# ---------------------------------------------------------------------------

def comb_step(frequency_hz: float, rate: int) -> int:
    """comb_kernels_reference.cpp:44 comb_step."""
    return int(_lib.lib().rvx_comb_step(frequency_hz, rate))


def comb_table() -> np.ndarray:
    """comb_kernels_reference.cpp:58 comb_table: 4096 complex64 (cos, sin)."""
    out = np.zeros(2 * 4096, dtype=np.float32)
    _lib.check(_lib.lib().rvx_comb_table(_fptr(out), out.size))
    return out.view(np.complex64)


def comb_check(params: _lib.CombParams) -> None:
    _lib.check(_lib.lib().rvx_comb_check(C.byref(params)))


# ---------------------------------------------------------------------------
# The per-receiver impulse blanker (core/dsp/noise_reference.h)
# ---------------------------------------------------------------------------

BLANK_DETECT, BLANK_APPLY = 0, 1


def design_blanker(channel_rate: int) -> _lib.BlankerConfig:
    """noise_reference.cpp:112 design_blanker."""
    out = _lib.BlankerConfig()
    _lib.check(_lib.lib().rvx_design_blanker(channel_rate, C.byref(out)))
    return out


def blanker_threshold(threshold_db: float = 12.0) -> float:
    return float(_lib.lib().rvx_blanker_threshold(threshold_db))


def blanker_check(config: _lib.BlankerConfig, params: _lib.BlankerParams, pass_: int) -> None:
    c = _lib.BlankerConfig.from_buffer_copy(bytes(config))
    c.pass_ = pass_
    _lib.check(_lib.lib().rvx_blanker_check(C.byref(c), C.byref(params)))


# ---------------------------------------------------------------------------
# The line kernel and spectral noise reduction (core/dsp/noise_reference.h)
# ---------------------------------------------------------------------------

LINE_NOTCH, LINE_ALE, LINE_GUARD = 1, 2, 4
LINE_NOTCH_COUNT_SHIFT = 8
MAX_NOTCHES = 4
MAX_LINE_STRIDE = 8
LINE_STATE_POSITION, LINE_STATE_GUARD, LINE_STATE_MORE, LINE_STATE_WEIGHTS = 4, 8, 12, 24
CW_GUARD_MIN_HZ, CW_GUARD_DEPTH_DB = 50.0, 40.0


def line_notch_flags(sections: int) -> int:
    """noise_reference.h:303 line_notch_flags."""
    return 0 if sections == 0 else LINE_NOTCH | ((sections - 1) << LINE_NOTCH_COUNT_SHIFT)


def line_notch_sections(flags: int) -> int:
    return 0 if not flags & LINE_NOTCH else 1 + ((flags >> LINE_NOTCH_COUNT_SHIFT) & 3)


def design_notch(audio_hz: float, depth_db: float, width_hz: float,
                 audio_rate: int) -> tuple[float, float, float, float, float]:
    """noise_reference.cpp:424 design_notch: (b0, b1, b2, a1, a2) in double."""
    out = (C.c_double * 5)()
    _lib.check(_lib.lib().rvx_design_notch(audio_hz, depth_db, width_hz, audio_rate, out))
    return tuple(out)


def notch_section_db(design, audio_hz: float, audio_rate: int) -> float:
    d = (C.c_double * 5)(*design)
    return float(_lib.lib().rvx_notch_section_db(d, audio_hz, audio_rate))


def notch_response_db(centre_hz, depth_db, width_hz, audio_rate, audio_hz) -> float:
    return notch_section_db(design_notch(centre_hz, depth_db, width_hz, audio_rate), audio_hz,
                            audio_rate)


def set_notch(params: _lib.LineParams, index: int, design, guard: bool = False) -> None:
    """Section `index` (0 the first, 1..3 the cascade) or the guard from a
    design, each coefficient rounded to float once."""
    b0, b1, b2, a1, a2 = (float(np.float32(v)) for v in design)
    if guard:
        params.gb0, params.gb1, params.gb2, params.ga1, params.ga2 = b0, b1, b2, a1, a2
    elif index == 0:
        params.b0, params.b1, params.b2, params.a1, params.a2 = b0, b1, b2, a1, a2
    else:
        params.more[index - 1] = _lib.LineSection(b0, b1, b2, a1, a2)


def line_state_size(config: _lib.LineConfig) -> int:
    return int(_lib.lib().rvx_line_state_size(C.byref(config)))


def line_check(config: _lib.LineConfig, params: _lib.LineParams) -> None:
    _lib.check(_lib.lib().rvx_line_check(C.byref(config), C.byref(params)))


def spectral_frame_for(audio_rate: int) -> int:
    """noise_reference.cpp:493 spectral_config_for."""
    return int(_lib.lib().rvx_spectral_frame_for(audio_rate))


def spectral_state_size(frame: int) -> int:
    return int(_lib.lib().rvx_spectral_state_size(frame))


def spectral_tables(frame: int) -> np.ndarray:
    """noise_reference.cpp:503 design_spectral_tables: cos, sin, analysis and
    synthesis (1/N folded in) windows, frame floats each."""
    out = np.zeros(4 * frame, np.float32)
    _lib.check(_lib.lib().rvx_design_spectral_tables(frame, out.ctypes.data_as(C.c_void_p),
                                                      out.size))
    return out


def spectral_params(strength: float) -> _lib.SpectralParams:
    out = _lib.SpectralParams()
    _lib.check(_lib.lib().rvx_spectral_params_for(strength, C.byref(out)))
    return out


def spectral_check(frame: int, params: _lib.SpectralParams) -> None:
    _lib.check(_lib.lib().rvx_spectral_check(frame, C.byref(params)))


# ---------------------------------------------------------------------------
# The display tap (core/dsp/vrx_reference.h:1639-1768)
# ---------------------------------------------------------------------------

DISPLAY_TAPS = 256


def display_max_decimation() -> int:
    return int(_lib.lib().rvx_display_max_decimation())


def display_ladder(channel_rate: int) -> list[int]:
    out = (C.c_uint32 * 64)()
    n = _lib.lib().rvx_display_ladder(channel_rate, out, 64)
    return list(out[:n])


def display_decimation(channel_rate: int, reach_hz: int) -> int:
    return int(_lib.lib().rvx_display_decimation(channel_rate, reach_hz))


@dataclass
class VrxDisplay:
    """A receiver's display tap: the plan block and its 256-tap table."""

    plan: _lib.VrxDisplayPlan
    taps: np.ndarray  # complex64, DISPLAY_TAPS + 1

    def block(self, channel_base: int, channel_mask: int, display_mask: int,
              first_output: int, count: int) -> tuple[_lib.VrxFineParams, int]:
        out = _lib.VrxFineParams()
        first_input = C.c_uint64()
        _lib.check(_lib.lib().rvx_display_block(C.byref(self.plan), channel_base, channel_mask,
                                                display_mask, first_output, count, C.byref(out),
                                                C.byref(first_input)))
        return out, first_input.value


def plan_vrx_display(plan: VrxPlan) -> VrxDisplay:
    """vrx_reference.cpp:2988 plan_vrx_display."""
    out = _lib.VrxDisplayPlan()
    taps = np.zeros(2 * (DISPLAY_TAPS + 1), np.float32)
    _lib.check(_lib.lib().rvx_plan_vrx_display(C.byref(plan.plan), C.byref(out),
                                               taps.ctypes.data_as(C.c_void_p), taps.size))
    return VrxDisplay(out, taps.view(np.complex64))

