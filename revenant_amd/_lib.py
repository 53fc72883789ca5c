"""ctypes binding of ``librvx_hip.so`` (the C ABI declared in include/rvx.h).

The library is required: importing the operators without it raises, and on a
machine with a GPU nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librvx_hip.so"

u32, i32, u64, i64, f32, f64, vp = (C.c_uint32, C.c_int32, C.c_uint64, C.c_int64, C.c_float,
                                    C.c_double, C.c_void_p)
fptr = C.POINTER(C.c_float)


class RvxError(RuntimeError):
    """A refused parameter block or a failed device call, with the library's message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("channels", u32), ("taps_per_branch", u32), ("decimation", u32)]


class ConvertParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("dst_offset", u32), ("src_offset", u32), ("count", u32)]


class PfbBranchParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("base_offset", u32), ("block_count", u32)]


class PfbFftParams(C.Structure):
    _fields_ = [("channels", u32), ("decimation", u32), ("stages", u32), ("block_base", u32),
                ("block_count", u32), ("out_ring_blocks", u32), ("out_ring_mask", u32)]


class SpectrumParams(C.Structure):
    _fields_ = [("channels", u32), ("transform", u32), ("stages", u32), ("chan_blocks", u32),
                ("chan_mask", u32), ("in_offset", u32)]


class SpectrumLevelsParams(C.Structure):
    _fields_ = [("bins", u32), ("low_permille", u32), ("high_permille", u32)]


class EmaParams(C.Structure):
    _fields_ = [("channels", u32), ("chan_blocks", u32), ("chan_mask", u32), ("in_offset", u32),
                ("frame_blocks", u32), ("frames", u32), ("alpha", f32), ("frame_decay", f32),
                ("row_frames", u32), ("row_decay", f32)]


class VrxFineConfig(C.Structure):
    _fields_ = [("taps", u32), ("phases", u32), ("nco_log2", u32)]


class VrxFineParams(C.Structure):
    _fields_ = [(n, u32) for n in ("chan_base", "chan_mask", "in_offset", "out_offset",
                                   "out_mask", "count", "step_whole", "step_rem", "out_rate",
                                   "frac0", "nco_phase_high", "nco_phase_low", "nco_delta_high",
                                   "nco_delta_low")] + [("inv_out_rate", f32)]


class VrxDemodConfig(C.Structure):
    _fields_ = [(n, u32) for n in ("mode", "decimation", "audio_taps", "dc_taps", "channels",
                                   "pilot_taps")]


class VrxDemodParams(C.Structure):
    _fields_ = [("in_mask", u32), ("in_offset", u32), ("count", u32), ("gain", f32)]


class IqMomentsParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("src_offset", u32), ("count", u32)]


class IqCorrectParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("offset", u32), ("count", u32), ("dc_i", f32),
                ("dc_q", f32), ("cross", f32), ("scale", f32)]


class IqMomentsTotal(C.Structure):
    _fields_ = [("samples", u64), ("sum_i", f64), ("sum_q", f64), ("sum_ii", f64),
                ("sum_qq", f64), ("sum_iq", f64)]


class FrontEndConfig(C.Structure):
    _fields_ = [("dc_time_constant_s", f64), ("iq_time_constant_s", f64),
                ("max_gain_error_db", f64), ("max_phase_error_deg", f64)]


class FrontEndEstimate(C.Structure):
    _fields_ = [("measured", i32), ("iq_plausible", i32), ("samples_seen", u64), ("dc_i", f64),
                ("dc_q", f64), ("gain", f64), ("sin_phase", f64)]


class ExcisionConfig(C.Structure):
    _fields_ = [(n, u32) for n in ("pass_", "seg_log2", "segments", "max_width", "hang", "lead")]


class ExcisionParams(C.Structure):
    _fields_ = [(n, u32) for n in ("ring_mask", "first", "count", "seg_mask", "flag_mask",
                                   "removed_mask")] + [("history", i32), ("threshold", f32)]


class ExcisionRings(C.Structure):
    _fields_ = [("seg_mask", u32), ("flag_mask", u32), ("removed_mask", u32)]


class ExcisionPlan(C.Structure):
    _fields_ = [("segment", ExcisionParams), ("reference", ExcisionParams),
                ("flag", ExcisionParams), ("apply", ExcisionParams), ("cleaned_end", u64),
                ("restarted", i32), ("reserved", u32)]


class CombLine(C.Structure):
    _fields_ = [("phase0", u32), ("step", u32), ("pad0", u32), ("pad1", u32),
                ("amplitude", f32 * 2)]


class CombEstimate(C.Structure):
    _fields_ = [("offset_hz", f64), ("spacing_hz", f64), ("k_lo", i32), ("k_hi", i32),
                ("lines_found", u32), ("median_height_db", f32), ("ripple_hz", f64),
                ("sideband_orders", i32), ("spur", u32), ("curvature_hz", f64)]


class CombStatus(C.Structure):
    _fields_ = [("lock", u32), ("lines_cancelled", u32), ("suppression_db", f32),
                ("spacing_drift_hz_per_s", f32), ("comb", CombEstimate), ("combs_active", u32),
                ("spurs_cancelled", u32), ("residual_rad", f32), ("reserved", u32)]


class SpurBand(C.Structure):
    _fields_ = [("low_hz", f64), ("high_hz", f64)]


COMB_SPUR_HOLD_OFF_S = 3.0  # RVX_COMB_SPUR_HOLD_OFF_S


class NoiseProfileSettings(C.Structure):
    _fields_ = [("enabled", i32), ("auto_learn", i32), ("strength_db", f64),
                ("learn_count", u32), ("reset_count", u32)]


NOISE_PROFILE_MAX_STRENGTH_DB = 20.0  # RVX_NOISE_PROFILE_MAX_STRENGTH_DB

COMB_IDLE, COMB_ACQUIRING, COMB_LOCKED, COMB_LOST, COMB_RECOVERING = 0, 1, 2, 3, 4
COMB_MIN_LINES, COMB_MAX_COMBS, COMB_MAX_LINES = 8, 8, 1024


class CombParams(C.Structure):
    _fields_ = [(n, u32) for n in ("ring_mask", "first", "count", "lines", "sub_blocks",
                                   "removed_mask", "pad0", "pad1")]


class BlankerConfig(C.Structure):
    _fields_ = [(n, u32) for n in ("pass_", "window", "guard", "hang", "lead")]


class BlankerParams(C.Structure):
    _fields_ = [(n, u32) for n in ("chan_base", "chan_mask", "chan_first", "out_mask",
                                   "out_first", "count")] + [("threshold", f32)]


class LineConfig(C.Structure):
    _fields_ = [("lanes", u32), ("taps", u32), ("history", u32)]

    def __init__(self, lanes=32, taps=128, history=2048):
        super().__init__(lanes, taps, history)


class LineSection(C.Structure):
    _fields_ = [(n, f32) for n in ("b0", "b1", "b2", "a1", "a2")]

    def __init__(self, b0=1.0, b1=0.0, b2=0.0, a1=0.0, a2=0.0):
        super().__init__(b0, b1, b2, a1, a2)


class LineParams(C.Structure):
    _fields_ = ([(n, u32) for n in ("count", "flags", "stride", "delay")] +
                [(n, f32) for n in ("b0", "b1", "b2", "a1", "a2", "mu", "leak", "eps",
                                    "gb0", "gb1", "gb2", "ga1", "ga2")] +
                [("more", LineSection * 3)])

    def __init__(self, **kw):
        super().__init__()
        defaults = dict(stride=1, delay=1, b0=1.0, leak=1.0, eps=1.0e-6, gb0=1.0)
        for k, v in {**defaults, **kw}.items():
            setattr(self, k, v)
        for i in range(3):
            if "more" not in kw:
                self.more[i] = LineSection()


class NoiseLineItem(C.Structure):
    _fields_ = [("params", LineParams), ("audio", vp), ("state", vp)]


class SpectralParams(C.Structure):
    _fields_ = [("count", u32)] + [(n, f32) for n in (
        "smoothing", "smoothing_complement", "gamma", "beta", "rise", "bias", "alpha",
        "floor_gain", "eps")]


class NoiseSpectralItem(C.Structure):
    _fields_ = [("params", SpectralParams), ("reserved", u32 * 2), ("audio", vp), ("state", vp)]


class VrxDisplayPlan(C.Structure):
    _fields_ = [("channel_rate", i64), ("display_rate", i64), ("decimation", u32),
                ("reserved", u32), ("reach_hz", i64), ("fine", VrxFineConfig),
                ("reserved2", u32), ("nco_delta", u64), ("pass_hz", f64), ("stop_hz", f64),
                ("attenuation_db", f64), ("group_delay_channel_samples", f64)]


class VrxCarrierConfig(C.Structure):
    _fields_ = [("power", u32), ("nco_log2", u32)]


class VrxCarrierParams(C.Structure):
    _fields_ = [(n, u32) for n in ("in_mask", "in_offset", "count", "flags")] + \
        [(n, f32) for n in ("shift", "k1_acquire", "k2_acquire", "k1_track", "k2_track",
                            "level_alpha", "lock_alpha", "clamp_radians")] + \
        [("hold_in", u32), ("hold_out", u32)]


class VrxCarrierItem(C.Structure):
    _fields_ = [("config", VrxCarrierConfig), ("params", VrxCarrierParams), ("fine_ring", vp),
                ("fine_samples", u64), ("state", vp), ("reserved", u64)]


class Carrier(C.Structure):
    _fields_ = [("kind", i32), ("reserved", i32), ("freq_hz", i64), ("tone_hz", i64),
                ("amplitude", f64), ("index", f64), ("depth", f64)]


class SynthParams(C.Structure):
    _fields_ = [("rate", i64), ("start", u64), ("count", u64), ("seed", u64),
                ("noise_sigma", f64), ("format", i32), ("reserved", i32)]


class ChainDesc(C.Structure):
    _fields_ = [("format", i32), ("channels", u32), ("taps_per_branch", u32),
                ("decimation", u32), ("input_origin", u64), ("input_samples", u64),
                ("first_block", u64), ("block_count", u32), ("out_ring_blocks", u32),
                ("frame_blocks", u32), ("alpha", f32), ("ring_rows", u32), ("taps_scaled", u32),
                ("channel_rows", vp), ("correction", vp), ("correction_before", vp),
                ("correction_from", u64), ("ring_pitch", u32), ("reserved", u32)]


RING_PITCH_PAD = 32  # RVX_RING_PITCH_PAD


class VrxRequest(C.Structure):
    _fields_ = [("center_hz", i64), ("bandwidth_hz", i64), ("passband_low", i64),
                ("passband_high", i64), ("mode", u32), ("deemphasis", u32), ("audio_rate", i64),
                ("cw_pitch", i64), ("mono", u32), ("reserved", u32)]


class VrxPlacement(C.Structure):
    _fields_ = [("channel", u32), ("bandwidth_clamped", u32), ("centre_num", i64),
                ("centre_den", i64), ("residual_num", i64), ("residual_den", i64),
                ("channel_rate", i64), ("granted_low", i64), ("granted_high", i64)]


class VrxPlan(C.Structure):
    _fields_ = [("placement", VrxPlacement), ("mode", u32), ("audio_decimation_taps", u32),
                ("channel_rate", i64), ("audio_rate", i64), ("demod_rate", i64),
                ("output_rate", i64), ("passband_low", i64), ("passband_high", i64),
                ("bandwidth", i64), ("deviation", i64), ("filter_num", i64), ("filter_den", i64),
                ("mix_num", i64), ("mix_den", i64), ("fine", VrxFineConfig), ("deemphasis", u32),
                ("fine_nco_delta", u64), ("demod", VrxDemodConfig), ("demod_gain", f32),
                ("audio_pass_hz", f64), ("audio_stop_hz", f64),
                ("carrier", VrxCarrierConfig), ("carrier_loop", VrxCarrierParams),
                ("stereo", u32)]


class VrxDesc(C.Structure):
    _fields_ = [("fine_config", VrxFineConfig), ("demod_config", VrxDemodConfig),
                ("fine", VrxFineParams), ("demod", VrxDemodParams), ("hist", u32), ("tile", u32),
                ("reserved", u32 * 2), ("channel_ring", vp), ("taps", vp), ("weights", vp),
                ("audio", vp)]


class VrxWork(C.Structure):
    _fields_ = [("vrx", u32), ("tile", u32)]


class VrxSpan(C.Structure):
    _fields_ = [("plan", C.POINTER(VrxPlan)), ("channel_base", u32), ("channel_mask", u32),
                ("channel_ring", vp), ("taps", vp), ("weights", vp), ("audio", vp),
                ("first_audio", u64), ("count", u32), ("tile", u32)]


class EngineConfig(C.Structure):
    _fields_ = [("format", i32), ("channels", u32), ("taps_per_branch", u32),
                ("decimation", u32), ("max_push", u32), ("ring_blocks", u32),
                ("frame_blocks", u32), ("row_frames", u32), ("alpha", f32), ("calibration", u32),
                ("start_index", u64), ("sample_rate", i64), ("ring_pitch", u32),
                ("reserved", u32)]


class EngineStats(C.Structure):
    _fields_ = [(n, u64) for n in ("samples_in", "next_block", "blocks", "spectrum_rows",
                                   "audio_lost", "audio_overrun", "grid_blocks")]


class EngineGrid(C.Structure):
    _fields_ = [(n, u32) for n in ("channels", "taps_per_branch", "decimation", "ring_blocks",
                                   "max_rows", "ring_pitch")]


class DetectConfig(C.Structure):
    _fields_ = [(n, u32) for n in ("bins", "noise_knots", "noise_window_knots", "max_sum_bins",
                                   "max_peaks", "split_gap_bins")] + \
               [(n, f64) for n in ("bin_width_hz", "detection_threshold_db", "noise_percentile",
                                   "noise_percentile_high", "noise_excision_sigma",
                                   "edge_floor_fraction", "edge_floor_sigma")]


class DetectBand(C.Structure):
    _fields_ = [("start", u32), ("width", u32), ("deflection", f64)]


class EventConfig(C.Structure):
    _fields_ = [("bin_width_hz", f64), ("bin_zero_hz", f64), ("source_center_hz", i64),
                ("source_rate", i64), ("detection_threshold_db", f64),
                ("occupied_power_fraction", f64), ("split_gap_db", f64), ("split_gap_bins", u32),
                ("residual_decisions", u32), ("average_seconds", f64),
                ("residual_rate_tolerance", f64), ("association_overlap", f64),
                ("confidence_rise", f64), ("confidence_half_life_seconds", f64),
                ("drop_confidence", f64), ("centre_smoothing", f64),
                ("bootstrap_hold_seconds", f64), ("birth_hits", u32), ("max_tracks", u32),
                ("grid_channels", u32), ("reserved", u32), ("channel_hysteresis", f64)]


class Candidate(C.Structure):
    _fields_ = [("first_bin", u32), ("last_bin", u32), ("center_hz", i64), ("bandwidth_hz", i64),
                ("snr_2500_db", f64), ("noise_floor_dbfs", f64), ("peak_dbfs", f64),
                ("margin_confidence", f64), ("shape_peak_to_mean", f64),
                ("shape_concentration", f64), ("shape_lower_fraction", f64),
                ("shape_skirt_fraction", f64), ("shape_skirt_bins_available", u32),
                ("shape_measured", u32)]


class Track(C.Structure):
    _fields_ = [("id", u64), ("state", u32), ("hits", u32), ("misses", u32),
                ("held_spells", u32), ("center_hz", i64), ("bandwidth_hz", i64),
                ("snr_2500_db", f64), ("confidence", f64), ("first_seen", u64),
                ("last_seen", u64), ("last_detected", u64), ("merged_into", u64),
                ("first_bin", u32), ("last_bin", u32), ("margin_confidence", f64),
                ("level_mean_db", f64), ("level_m2", f64), ("level_first_bin", u32),
                ("level_last_bin", u32), ("level_frames", u32), ("level_bins_known", u32),
                ("channel_index", i64), ("channel", u32), ("channel_valid", u32),
                ("shape_peak_to_mean", f64), ("shape_concentration", f64),
                ("shape_lower_fraction", f64), ("shape_skirt_fraction", f64),
                ("shape_skirt_bins_available", u32), ("shape_measured", u32),
                ("classification", u32), ("double_sideband", u32), ("keyups", u32),
                ("spur", u32), ("lo_spur", u32), ("reserved_spur", u32), ("emitter", u64),
                ("hd_lower_dbc", f64), ("hd_upper_dbc", f64), ("hd_passes", u32),
                ("hd_fails", u32), ("hd_hybrid", u32), ("reserved_hd", u32)]


class HdReading(C.Structure):
    _fields_ = [("measured", u32), ("hybrid", u32), ("centre_hz", f64), ("lower_dbc", f64),
                ("upper_dbc", f64), ("lower_over_floor_db", f64), ("upper_over_floor_db", f64),
                ("lower_over_gap_db", f64), ("upper_over_gap_db", f64), ("lower_tilt_db", f64),
                ("upper_tilt_db", f64)]


FAMILY_UNKNOWN, FAMILY_UNMODULATED, FAMILY_OTHER = 0, 1, 2
SPUR_BOOK_CAPACITY = 256


class SpurSighting(C.Structure):
    _fields_ = [("baseband_hz", i64), ("center_hz", i64), ("last_seen", u64)]


class TrackerStats(C.Structure):
    _fields_ = [(n, u64) for n in ("decisions", "candidates", "candidates_residual",
                                   "residual_parts_degenerate", "tracks_born", "tracks_dropped",
                                   "births_refused", "merges", "splits")]


VRX_TILE_SPAN = 3072  # RVX_VRX_TILE_SPAN
PREC_BF16X3, PREC_FP16 = 1, 2  # RVX_PREC_*

assert C.sizeof(VrxDesc) == 160 and C.sizeof(VrxWork) == 8 and C.sizeof(VrxSpan) == 64
assert C.sizeof(VrxFineParams) == 60 and C.sizeof(VrxDemodParams) == 16
assert C.sizeof(ConvertParams) == 16 and C.sizeof(PfbBranchParams) == 12

_SIGNATURES = {
    "rvx_abi_version": (i32, []),
    "rvx_last_error": (C.c_char_p, []),
    "rvx_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "rvx_ctx_destroy": (None, [vp]),
    "rvx_ctx_set_stream": (i32, [vp, vp]),
    "rvx_ctx_use_own_stream": (i32, [vp]),
    "rvx_ctx_stream": (vp, [vp]),
    "rvx_stream_sync": (i32, [vp]),
    "rvx_buf_alloc": (i32, [vp, u64, i32, C.POINTER(vp)]),
    "rvx_buf_free": (i32, [vp, vp]),
    "rvx_buf_write": (i32, [vp, vp, vp, u64]),
    "rvx_buf_read": (i32, [vp, vp, vp, u64]),
    "rvx_buf_clear": (i32, [vp, vp, u64]),
    "rvx_timer_begin": (i32, [vp]),
    "rvx_timer_end": (i32, [vp, fptr]),
    "rvx_design_prototype": (i32, [C.POINTER(Grid), f64, f64, fptr, u64, C.POINTER(f64),
                                   C.POINTER(u64)]),
    "rvx_build_twiddles": (i32, [u32, fptr, u64]),
    "rvx_build_spectrum_window": (i32, [u32, u32, f64, fptr, u64]),
    "rvx_build_nco_table": (i32, [u32, fptr, u64]),
    "rvx_turn_fixed64": (u64, [i64, i64]),
    "rvx_nco_delta": (i32, [i64, i64, i64, C.POINTER(u64)]),
    "rvx_convert": (i32, [vp, i32, vp, u64, C.POINTER(ConvertParams), vp, u64]),
    "rvx_cmul": (i32, [vp, vp, vp, vp, u64]),
    "rvx_pfb_branch": (i32, [vp, C.POINTER(Grid), vp, u64, vp, u64, C.POINTER(PfbBranchParams),
                             vp, u64]),
    "rvx_pfb_fft": (i32, [vp, C.POINTER(PfbFftParams), vp, u64, vp, u64, vp, u64]),
    "rvx_fft_radix2": (i32, [vp, vp, u64, vp, vp, u32, u32]),
    "rvx_spectrum": (i32, [vp, C.POINTER(SpectrumParams), vp, u64, vp, u64, vp, u64, vp, u64]),
    "rvx_spectrum_levels": (i32, [vp, C.POINTER(SpectrumLevelsParams), vp, u64, vp]),
    "rvx_ema_partial": (i32, [vp, C.POINTER(EmaParams), vp, u64, vp, u64]),
    "rvx_ema_combine": (i32, [vp, C.POINTER(EmaParams), vp, u64, vp, vp, u64]),
    "rvx_vrx_fine": (i32, [vp, C.POINTER(VrxFineConfig), C.POINTER(VrxFineParams), vp, u64, vp,
                           u64, vp, u64, vp, u64]),
    "rvx_vrx_demod": (i32, [vp, C.POINTER(VrxDemodConfig), C.POINTER(VrxDemodParams), vp, u64,
                            vp, u64, vp, u64]),
    "rvx_place": (i32, [C.POINTER(Grid), i64, C.POINTER(VrxRequest), C.POINTER(VrxPlacement)]),
    "rvx_channel_count_for": (u32, [i64, i64]),
    "rvx_resolve_passband": (i32, [C.POINTER(VrxRequest), C.POINTER(i64), C.POINTER(i64)]),
    "rvx_default_passband": (i32, [u32, C.POINTER(i64), C.POINTER(i64)]),
    "rvx_max_channel_bandwidth": (i64, [C.POINTER(VrxPlacement)]),
    "rvx_clamp_to_channel": (i32, [C.POINTER(VrxPlacement), C.POINTER(i64), C.POINTER(i64)]),
    "rvx_minimum_demod_rate": (i64, [u32, i64, i64]),
    "rvx_demod_rate_for": (i32, [C.POINTER(VrxRequest), C.POINTER(VrxPlacement), i64,
                                 C.POINTER(i64)]),
    "rvx_channel_centre": (i32, [C.POINTER(Grid), i64, u32, C.POINTER(i64), C.POINTER(i64)]),
    "rvx_channel_sample_to_input": (u64, [C.POINTER(Grid), u64, u64, u64]),
    "rvx_plan_vrx": (i32, [C.POINTER(Grid), i64, C.POINTER(VrxRequest), C.POINTER(VrxPlacement),
                           C.POINTER(VrxPlan)]),
    "rvx_plan_vrx_tables": (i32, [C.POINTER(VrxPlan), fptr, u64, fptr, u64]),
    "rvx_fine_block": (i32, [C.POINTER(VrxPlan), u32, u32, u32, u64, u32,
                             C.POINTER(VrxFineParams), C.POINTER(u64)]),
    "rvx_demod_block": (i32, [C.POINTER(VrxPlan), u32, u64, u32, C.POINTER(VrxDemodParams)]),
    "rvx_demod_history": (u32, [C.POINTER(VrxDemodConfig)]),
    "rvx_carrier_power": (u32, [u32]),
    "rvx_carrier_loop_gains": (i32, [f64, f64, i64, C.POINTER(f64), C.POINTER(f64)]),
    "rvx_design_carrier_loop": (i32, [i64, i64, C.POINTER(VrxCarrierParams)]),
    "rvx_carrier_retune": (i32, [i64, i64, i64, C.POINTER(i32), C.POINTER(f64)]),
    "rvx_carrier_block": (i32, [C.POINTER(VrxPlan), u32, u64, u32, i32, f32,
                                C.POINTER(VrxCarrierParams)]),
    "rvx_vrx_carrier_check": (i32, [C.POINTER(VrxCarrierConfig), C.POINTER(VrxCarrierParams)]),
    "rvx_vrx_carrier_item_check": (i32, [C.POINTER(VrxCarrierItem), u64]),
    "rvx_vrx_carrier": (i32, [vp, C.POINTER(VrxCarrierConfig), C.POINTER(VrxCarrierParams), vp,
                              u64, vp, u64, vp]),
    "rvx_vrx_carrier_batch": (i32, [vp, vp, u32, vp, u64]),
    "rvx_design_excision": (i32, [i64, f64, C.POINTER(ExcisionConfig)]),
    "rvx_excision_threshold": (f32, [f64]),
    "rvx_excision_check": (i32, [C.POINTER(ExcisionConfig), C.POINTER(ExcisionParams)]),
    "rvx_size_excision_rings": (i32, [C.POINTER(ExcisionConfig), u64, u32, u32,
                                      C.POINTER(ExcisionRings)]),
    "rvx_excision_cursor_create": (i32, [C.POINTER(ExcisionConfig), u32, C.POINTER(ExcisionRings),
                                         f32, C.POINTER(vp)]),
    "rvx_excision_cursor_destroy": (None, [vp]),
    "rvx_excision_cursor_reset": (None, [vp]),
    "rvx_excision_cursor_plan": (i32, [vp, u64, u32, C.POINTER(ExcisionPlan)]),
    "rvx_excise_segment": (i32, [vp, C.POINTER(ExcisionConfig), C.POINTER(ExcisionParams), vp, u64,
                                 vp, u64]),
    "rvx_excise_reference": (i32, [vp, C.POINTER(ExcisionConfig), C.POINTER(ExcisionParams), vp,
                                   vp, u64]),
    "rvx_excise_flag": (i32, [vp, C.POINTER(ExcisionConfig), C.POINTER(ExcisionParams), vp, u64,
                              vp, u64, vp, u64]),
    "rvx_excise_apply": (i32, [vp, C.POINTER(ExcisionConfig), C.POINTER(ExcisionParams), vp, u64,
                               vp, u64, vp, u64, vp, u64, vp]),
    "rvx_design_blanker": (i32, [i64, C.POINTER(BlankerConfig)]),
    "rvx_blanker_threshold": (f32, [f64]),
    "rvx_blanker_check": (i32, [C.POINTER(BlankerConfig), C.POINTER(BlankerParams)]),
    "rvx_blank_detect": (i32, [vp, C.POINTER(BlankerConfig), C.POINTER(BlankerParams), vp, u64,
                               vp, u64]),
    "rvx_blank_apply": (i32, [vp, C.POINTER(BlankerConfig), C.POINTER(BlankerParams), vp, u64,
                              vp, u64, vp, u64]),
    "rvx_design_notch": (i32, [f64, f64, f64, i64, vp]),
    "rvx_notch_section_db": (f64, [vp, f64, i64]),
    "rvx_line_state_size": (u64, [C.POINTER(LineConfig)]),
    "rvx_line_check": (i32, [C.POINTER(LineConfig), C.POINTER(LineParams)]),
    "rvx_noise_line_item_check": (i32, [C.POINTER(LineConfig), C.POINTER(NoiseLineItem), u64,
                                        u64]),
    "rvx_noise_line": (i32, [vp, C.POINTER(LineConfig), C.POINTER(LineParams), vp, u64, vp,
                             u64]),
    "rvx_noise_line_batch": (i32, [vp, C.POINTER(LineConfig), vp, u32]),
    "rvx_spectral_frame_for": (u32, [i64]),
    "rvx_spectral_state_size": (u64, [u32]),
    "rvx_design_spectral_tables": (i32, [u32, vp, u64]),
    "rvx_spectral_params_for": (i32, [f64, C.POINTER(SpectralParams)]),
    "rvx_spectral_check": (i32, [u32, C.POINTER(SpectralParams)]),
    "rvx_noise_spectral_item_check": (i32, [u32, C.POINTER(NoiseSpectralItem), u64, u64]),
    "rvx_noise_spectral": (i32, [vp, u32, C.POINTER(SpectralParams), vp, u64, vp, u64, vp, u64]),
    "rvx_noise_spectral_batch": (i32, [vp, u32, vp, u64, vp, u32]),
    "rvx_display_max_decimation": (u32, []),
    "rvx_display_ladder": (u32, [i64, vp, u32]),
    "rvx_display_decimation": (u32, [i64, i64]),
    "rvx_plan_vrx_display": (i32, [C.POINTER(VrxPlan), C.POINTER(VrxDisplayPlan), vp, u64]),
    "rvx_display_block": (i32, [C.POINTER(VrxDisplayPlan), u32, u32, u32, u64, u32,
                                C.POINTER(VrxFineParams), C.POINTER(u64)]),
    "rvx_engine_set_display": (i32, [vp, u32, u32, u32]),
    "rvx_engine_passband": (i32, [vp, u32, vp, u64, C.POINTER(u64), C.POINTER(i64)]),
    "rvx_engine_set_blanker": (i32, [vp, u32, C.POINTER(BlankerConfig), f32]),
    "rvx_engine_set_noise": (i32, [vp, u32, C.POINTER(LineConfig), C.POINTER(LineParams), u32,
                                   C.POINTER(SpectralParams)]),
    "rvx_design_deemphasis": (i32, [u32, i64, vp, u32, C.POINTER(u32)]),
    "rvx_comb_step": (u32, [f64, i64]),
    "rvx_comb_detect": (i32, [fptr, u64, i64, C.POINTER(CombEstimate), u32, u32,
                              C.POINTER(CombEstimate), u32, C.POINTER(u32)]),
    "rvx_noise_profile_create": (i32, [C.POINTER(vp)]),
    "rvx_noise_profile_destroy": (None, [vp]),
    "rvx_noise_profile_process": (i32, [vp, fptr, u64, f64, i32,
                                        C.POINTER(NoiseProfileSettings)]),
    "rvx_noise_profile_retune": (i32, [vp]),
    "rvx_noise_profile_state": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(u32),
                                      C.POINTER(u32)]),
    "rvx_noise_profile_estimate": (i32, [vp, C.POINTER(f64), u32, C.POINTER(u32)]),
    "rvx_comb_bank_create": (i32, [i64, C.POINTER(vp)]),
    "rvx_comb_bank_destroy": (None, [vp]),
    "rvx_comb_bank_reset": (i32, [vp]),
    "rvx_comb_bank_detected": (i32, [vp, C.POINTER(CombEstimate), u32]),
    "rvx_comb_bank_spurs": (i32, [vp, C.POINTER(CombEstimate), u32]),
    "rvx_comb_bank_drop_spurs": (i32, [vp]),
    "rvx_comb_bank_withdraw_spurs": (i32, [vp, C.POINTER(f64), u32]),
    "rvx_comb_bank_drop_spurs_in": (i32, [vp, C.POINTER(SpurBand), u32]),
    "rvx_comb_bank_plan": (i32, [vp, u64, u32, C.POINTER(CombLine), u32, C.POINTER(u32),
                                 C.POINTER(u32), C.POINTER(i32)]),
    "rvx_comb_bank_observe": (i32, [vp, u64, u32, u32, fptr, u64]),
    "rvx_comb_bank_status": (i32, [vp, C.POINTER(CombStatus)]),
    "rvx_comb_bank_active": (i32, [vp, C.POINTER(CombEstimate), u32, C.POINTER(u32)]),
    "rvx_comb_table": (i32, [fptr, u64]),
    "rvx_comb_check": (i32, [C.POINTER(CombParams)]),
    "rvx_comb_project": (i32, [vp, C.POINTER(CombParams), vp, u64, vp, u64, vp, u64, vp, u64]),
    "rvx_comb_subtract": (i32, [vp, C.POINTER(CombParams), vp, u64, vp, u64, vp, u64, vp, u64,
                                vp, u64]),
    "rvx_iq_moments": (i32, [vp, C.POINTER(IqMomentsParams), vp, u64, vp, u64]),
    "rvx_iq_moments_codes": (i32, [vp, i32, vp, u64, u64, u32, vp, u64]),
    "rvx_iq_correct": (i32, [vp, C.POINTER(IqCorrectParams), vp, u64, vp, u64]),
    "rvx_total_moments": (i32, [fptr, u64, u32, C.POINTER(IqMomentsTotal)]),
    "rvx_front_end_default_config": (None, [C.POINTER(FrontEndConfig)]),
    "rvx_front_end_create": (i32, [C.POINTER(FrontEndConfig), C.POINTER(vp)]),
    "rvx_front_end_destroy": (None, [vp]),
    "rvx_front_end_reset": (None, [vp]),
    "rvx_front_end_update": (i32, [vp, C.POINTER(IqMomentsTotal), i64]),
    "rvx_front_end_estimate_get": (i32, [vp, C.POINTER(FrontEndEstimate)]),
    "rvx_front_end_correction": (i32, [vp, i32, i32, C.POINTER(IqCorrectParams)]),
    "rvx_front_end_derived": (i32, [C.POINTER(FrontEndEstimate), C.POINTER(f64), C.POINTER(f64),
                                    C.POINTER(f64), C.POINTER(f64)]),
    "rvx_synth": (i32, [vp, C.POINTER(Carrier), u32, C.POINTER(SynthParams), vp, u64]),
    "rvx_chain": (i32, [vp, C.POINTER(ChainDesc), vp, vp, u64, vp, u64, vp, u64, vp, u64]),
    "rvx_vrx_batch": (i32, [vp, vp, u32, vp, u32, vp, vp, u64, u32]),
    "rvx_vrx_batch_noalloc": (i32, [vp, vp, u32, vp, u32, vp, u64, u32]),
    "rvx_vrx_desc_check": (i32, [C.POINTER(VrxDesc)]),
    "rvx_vrx_batch_plan": (i32, [C.POINTER(VrxSpan), u32, C.POINTER(VrxDesc), u32,
                                 C.POINTER(u32), C.POINTER(VrxWork), u32, C.POINTER(u32)]),
    "rvx_engine_create": (i32, [vp, C.POINTER(EngineConfig), fptr, u64, C.POINTER(vp)]),
    "rvx_engine_destroy": (None, [vp]),
    "rvx_engine_add_vrx": (i32, [vp, C.POINTER(VrxPlan), fptr, u64, fptr, u64, u32, u32,
                                 C.POINTER(u32)]),
    "rvx_engine_add_vrx_on": (i32, [vp, u32, C.POINTER(VrxPlan), fptr, u64, fptr, u64, u32, u32,
                                    C.POINTER(u32)]),
    "rvx_engine_add_grid": (i32, [vp, C.POINTER(EngineGrid), fptr, u64, C.POINTER(u32)]),
    "rvx_engine_grid_ring": (vp, [vp, u32]),
    "rvx_engine_set_excision": (i32, [vp, C.POINTER(ExcisionConfig), f32]),
    "rvx_engine_excision_counters": (i32, [vp, C.POINTER(u32)]),
    "rvx_engine_set_comb": (i32, [vp, C.POINTER(CombLine), u32, u64, i32]),
    "rvx_engine_comb_projections": (i32, [vp, fptr, u64, C.POINTER(u64), C.POINTER(u64),
                                          C.POINTER(u32), C.POINTER(u32)]),
    "rvx_engine_grid_row": (i32, [vp, u32, u32, C.POINTER(u32), C.POINTER(u64)]),
    "rvx_engine_remove_vrx": (i32, [vp, u32]),
    "rvx_engine_push": (i32, [vp, vp, u64, i32]),
    "rvx_engine_read_audio": (i32, [vp, u32, fptr, u64, C.POINTER(u64), C.POINTER(u64)]),
    "rvx_engine_spectrum": (i32, [vp, fptr, u64]),
    "rvx_engine_channel_ring": (vp, [vp]),
    "rvx_engine_stats_get": (i32, [vp, C.POINTER(EngineStats)]),
    "rvx_engine_front_end": (i32, [vp, C.POINTER(FrontEndEstimate), C.POINTER(IqCorrectParams)]),
    "rvx_detector_create": (i32, [vp, C.POINTER(DetectConfig), C.POINTER(vp)]),
    "rvx_detector_destroy": (None, [vp]),
    "rvx_event_config_defaults": (None, [C.POINTER(EventConfig)]),
    "rvx_candidates": (i32, [C.POINTER(EventConfig), vp, vp, u32, C.POINTER(DetectBand), u32,
                             C.POINTER(Candidate), u32, C.POINTER(u32)]),
    "rvx_tracker_create": (i32, [C.POINTER(EventConfig), u32, C.POINTER(vp)]),
    "rvx_tracker_destroy": (None, [vp]),
    "rvx_tracker_measure_levels": (i32, [vp, fptr, u32]),
    "rvx_tracker_update": (i32, [vp, vp, vp, C.POINTER(Candidate), u32, u64, f64,
                                 C.POINTER(u32)]),
    "rvx_tracker_tracks": (i32, [vp, C.POINTER(Track), u32, C.POINTER(u32)]),
    "rvx_tracker_stats_get": (i32, [vp, C.POINTER(TrackerStats)]),
    "rvx_tracker_record_probe": (i32, [vp, u64, u32, u32]),
    "rvx_measure_hd_sidebands": (i32, [vp, vp, u32, f64, f64, f64, C.POINTER(HdReading)]),
    "rvx_tracker_withdrawn_spurs": (i32, [vp, C.POINTER(i64), u32, C.POINTER(u32)]),
    "rvx_tracker_spur_book": (i32, [vp, C.POINTER(SpurSighting), u32, C.POINTER(u32)]),
    "rvx_tracker_adopt_spur_book": (i32, [vp, C.POINTER(SpurSighting), u32]),
    "rvx_detect": (i32, [vp, vp, vp, C.POINTER(DetectBand), u32, C.POINTER(u32),
                         C.POINTER(u32)]),
    "rvx_detector_floor": (vp, [vp]),
    "rvx_detector_sigma": (vp, [vp]),
    "rvx_net_create": (i32, [vp, u32, u32, u32, u64, C.POINTER(vp)]),
    "rvx_net_destroy": (None, [vp]),
    "rvx_net_param_count": (u32, [vp]),
    "rvx_net_params": (i32, [vp, fptr, u64, i32]),
    "rvx_net_grads": (i32, [vp, fptr, u64]),
    "rvx_net_forward": (i32, [vp, vp, u32, vp]),
    "rvx_net_train_step": (i32, [vp, vp, vp, u32, f32, f32, C.POINTER(f32)]),
    "rvx_to_bf16": (i32, [vp, vp, u64, vp]),
    "rvx_chain_mfma": (i32, [vp, C.POINTER(ChainDesc), i32, vp, vp, u64, vp, u64, vp, u64, vp,
                             u64]),
    "rvx_spectrum_fp8": (i32, [vp, vp, u64, vp]),
    "rvx_prescale_taps": (i32, [fptr, u64, fptr]),
    "rvx_graph_begin": (i32, [vp]),
    "rvx_graph_end": (i32, [vp, C.POINTER(vp)]),
    "rvx_graph_launch": (i32, [vp, vp]),
    "rvx_graph_destroy": (None, [vp]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m revenant_amd.build` (there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    """Names the C ABI promises (every function declared in include/rvx.h)."""
    return list(_SIGNATURES)


def check(code: int) -> None:
    if code != 0:
        msg = lib().rvx_last_error()
        raise RvxError(code, msg.decode() if msg else "")
