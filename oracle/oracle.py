"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, the smoke check and the
benchmark's CPU-baseline leg. It is the checker, never the thing measured or
shipped.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"

u32, i32, u64, f32, vp = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_void_p


class OChain(C.Structure):
    _fields_ = [("format", i32), ("channels", u32), ("taps_per_branch", u32),
                ("decimation", u32), ("first_block", u64), ("block_count", u32),
                ("out_ring_blocks", u32), ("frame_blocks", u32), ("alpha", f32),
                ("frame_decay", f32)]


class ODetectConfig(C.Structure):
    _fields_ = [(n, u32) for n in ("bins", "noise_knots", "noise_window_knots", "max_sum_bins",
                                   "max_peaks", "split_gap_bins")] + \
               [(n, C.c_double) for n in ("bin_width_hz", "detection_threshold_db",
                                          "noise_percentile", "noise_percentile_high",
                                          "noise_excision_sigma", "edge_floor_fraction",
                                          "edge_floor_sigma")]


class OBand(C.Structure):
    _fields_ = [("start", u32), ("width", u32), ("deflection", C.c_double)]


def detect_config(bins: int, bin_width_hz: float, threshold_db: float = 6.0, **kw) -> ODetectConfig:
    """DetectorConfig's defaults (core/detect/detector.h:579-909)."""
    c = ODetectConfig(bins, kw.get("noise_knots", 32), kw.get("noise_window_knots", 8),
                      kw.get("max_sum_bins", 0), kw.get("max_peaks", 0),
                      kw.get("split_gap_bins", 8), bin_width_hz, threshold_db,
                      kw.get("noise_percentile", 0.10), kw.get("noise_percentile_high", 0.40),
                      kw.get("noise_excision_sigma", 4.0), kw.get("edge_floor_fraction", 0.0),
                      kw.get("edge_floor_sigma", 3.0))
    return c


class OEventConfig(C.Structure):
    _fields_ = [("bin_width_hz", C.c_double), ("bin_zero_hz", C.c_double),
                ("source_center_hz", C.c_int64), ("source_rate", C.c_int64),
                ("detection_threshold_db", C.c_double), ("occupied_power_fraction", C.c_double),
                ("split_gap_db", C.c_double), ("split_gap_bins", u32),
                ("residual_decisions", u32), ("average_seconds", C.c_double),
                ("residual_rate_tolerance", C.c_double), ("association_overlap", C.c_double),
                ("confidence_rise", C.c_double), ("confidence_half_life_seconds", C.c_double),
                ("drop_confidence", C.c_double), ("centre_smoothing", C.c_double),
                ("bootstrap_hold_seconds", C.c_double), ("birth_hits", u32),
                ("max_tracks", u32)]


class OCandidate(C.Structure):
    _fields_ = [("first_bin", u32), ("last_bin", u32), ("center", C.c_int64),
                ("bandwidth", C.c_int64), ("snr_2500_db", C.c_double),
                ("noise_floor_dbfs", C.c_double), ("peak_dbfs", C.c_double)]


class OTrack(C.Structure):
    _fields_ = [("id", u64), ("state", u32), ("hits", u32), ("misses", u32),
                ("held_spells", u32), ("center", C.c_int64), ("bandwidth", C.c_int64),
                ("snr_2500_db", C.c_double), ("confidence", C.c_double), ("first_seen", u64),
                ("last_seen", u64), ("last_detected", u64), ("merged_into", u64),
                ("first_bin", u32), ("last_bin", u32)]


def event_config(bin_width_hz: float, bin_zero_hz: float, source_rate: int,
                 source_center_hz: int = 0, **kw) -> OEventConfig:
    """DetectorConfig's defaults for the event stages (core/detect/detector.h:600-1200)."""
    d = dict(detection_threshold_db=6.0, occupied_power_fraction=0.99, split_gap_db=1.0,
             split_gap_bins=8, residual_decisions=4, average_seconds=1.0,
             residual_rate_tolerance=0.25, association_overlap=0.3, confidence_rise=0.35,
             confidence_half_life_seconds=1.0, drop_confidence=0.05, centre_smoothing=0.4,
             bootstrap_hold_seconds=3.0, birth_hits=3, max_tracks=512)
    d.update(kw)
    c = OEventConfig()
    c.bin_width_hz, c.bin_zero_hz = bin_width_hz, bin_zero_hz
    c.source_center_hz, c.source_rate = source_center_hz, source_rate
    for k, v in d.items():
        setattr(c, k, v)
    return c


def candidates(cfg: OEventConfig, average: np.ndarray, floor: np.ndarray, bands) -> list:
    avg = np.ascontiguousarray(average, dtype=np.float64)
    flo = np.ascontiguousarray(floor, dtype=np.float64)
    arr = (OBand * max(1, len(bands)))(*[OBand(int(s), int(w), float(d)) for s, w, d in bands])
    cap = 4 * len(bands) + 4
    out = (OCandidate * cap)()
    n = u32()
    _check(lib().oracle_candidates(C.byref(cfg), _a(avg), _a(flo), u32(avg.size), arr,
                                   u32(len(bands)), out, u32(cap), C.byref(n)), "candidates")
    return [out[i] for i in range(n.value)]


class OracleTracker:
    """reject_residual + update_tracks, decision by decision."""

    def __init__(self, cfg: OEventConfig, bins: int):
        lib().oracle_tracker_create.restype = C.c_void_p
        self.h = C.c_void_p(lib().oracle_tracker_create(C.byref(cfg), u32(bins)))

    def update(self, average, floor, cands, now: int, elapsed: float) -> int:
        avg = np.ascontiguousarray(average, dtype=np.float64)
        flo = np.ascontiguousarray(floor, dtype=np.float64)
        arr = (OCandidate * max(1, len(cands)))(*cands)
        kept = u32()
        _check(lib().oracle_tracker_update(self.h, _a(avg), _a(flo), arr, u32(len(cands)),
                                           u64(now), C.c_double(elapsed), C.byref(kept)),
               "tracker_update")
        return kept.value

    def tracks(self) -> list:
        n = u32()
        cap = 4096
        out = (OTrack * cap)()
        _check(lib().oracle_tracker_tracks(self.h, out, u32(cap), C.byref(n)), "tracker_tracks")
        return [out[i] for i in range(n.value)]

    def close(self) -> None:
        if self.h:
            lib().oracle_tracker_destroy(self.h)
            self.h = None


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB_PATH))
        for name in ("oracle_convert", "oracle_cmul", "oracle_pfb_branch", "oracle_pfb_fft",
                     "oracle_fft_radix2", "oracle_spectrum", "oracle_spectrum_levels",
                     "oracle_det", "oracle_ema_partial", "oracle_ema_combine", "oracle_chain",
                     "fp64_channelize", "fp64_dft", "oracle_vrx_fine", "oracle_vrx_demod",
                     "oracle_vrx_carrier", "oracle_iq_moments", "oracle_iq_correct",
                     "oracle_excise_segment", "oracle_excise_reference", "oracle_excise_flag",
                     "oracle_excise_apply", "oracle_comb_project", "oracle_comb_subtract",
                     "oracle_blank_detect", "oracle_blank_apply", "oracle_noise_line",
                     "oracle_noise_spectral"):
            getattr(_lib, name).restype = i32
    return _lib


def _a(x: np.ndarray):
    return x.ctypes.data_as(vp)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ValueError(f"oracle refused {what} (rc={rc})")


def f32c(x: np.ndarray) -> np.ndarray:
    """Complex samples as contiguous float32 pairs."""
    if np.iscomplexobj(x):
        x = np.ascontiguousarray(x.astype(np.complex64)).view(np.float32)
    return np.ascontiguousarray(x, dtype=np.float32)


def convert(fmt: int, packed: np.ndarray, params, ring: np.ndarray) -> None:
    """ring (float32 pairs) updated in place. params: ctypes ConvertParams."""
    _check(lib().oracle_convert(i32(fmt), _a(packed), u64(packed.nbytes), C.byref(params),
                                _a(ring), u64(ring.size // 2)), "convert")


def cmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a, b = f32c(a), f32c(b)
    out = np.empty_like(a)
    _check(lib().oracle_cmul(_a(a), _a(b), _a(out), u64(a.size // 2)), "cmul")
    return out


def pfb_branch(grid, taps: np.ndarray, ring: np.ndarray, params) -> np.ndarray:
    taps = np.ascontiguousarray(taps, dtype=np.float32)
    ring = f32c(ring)
    out = np.zeros(2 * params.block_count * grid.channels, dtype=np.float32)
    _check(lib().oracle_pfb_branch(C.byref(grid), _a(taps), u64(taps.size), _a(ring),
                                   u64(ring.size // 2), C.byref(params), _a(out),
                                   u64(out.size // 2)), "pfb_branch")
    return out


def pfb_fft(params, branch: np.ndarray, twiddles: np.ndarray, ring: np.ndarray) -> None:
    branch, tw = f32c(branch), f32c(twiddles)
    _check(lib().oracle_pfb_fft(C.byref(params), _a(branch), u64(branch.size // 2), _a(tw),
                                u64(tw.size // 2), _a(ring), u64(ring.size // 2)), "pfb_fft")


def fft_radix2(twiddles: np.ndarray, x: np.ndarray, size: int, batch: int) -> np.ndarray:
    tw, x = f32c(twiddles), f32c(x)
    out = np.zeros(2 * size * batch, dtype=np.float32)
    _check(lib().oracle_fft_radix2(_a(tw), u64(tw.size // 2), _a(x), _a(out), u32(size),
                                   u32(batch)), "fft_radix2")
    return out


def spectrum(params, ring: np.ndarray, twiddles: np.ndarray, window: np.ndarray) -> np.ndarray:
    ring, tw = f32c(ring), f32c(twiddles)
    window = np.ascontiguousarray(window, dtype=np.float32)
    frame = np.zeros(params.channels * (params.transform // 2), dtype=np.float32)
    _check(lib().oracle_spectrum(C.byref(params), _a(ring), u64(ring.size // 2), _a(tw),
                                 u64(tw.size // 2), _a(window), u64(window.size), _a(frame),
                                 u64(frame.size)), "spectrum")
    return frame


def spectrum_levels(params, frame: np.ndarray) -> np.ndarray:
    frame = np.ascontiguousarray(frame, dtype=np.float32)
    out = np.zeros(2, dtype=np.float32)
    _check(lib().oracle_spectrum_levels(C.byref(params), _a(frame), u64(frame.size), _a(out)),
           "spectrum_levels")
    return out


DET = {"log2": 0, "recip": 1, "rsqrt": 2, "sqrt": 3, "atan2": 4, "atan_unit": 5}


def det(which: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a if b is None else np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros_like(a)
    _check(lib().oracle_det(i32(DET[which]), _a(a), _a(b), _a(out), u64(a.size)), which)
    return out


def ema_partial(params, ring: np.ndarray) -> np.ndarray:
    ring = f32c(ring)
    out = np.zeros(params.frames * params.channels, dtype=np.float32)
    _check(lib().oracle_ema_partial(C.byref(params), _a(ring), u64(ring.size // 2), _a(out),
                                    u64(out.size)), "ema_partial")
    return out


def ema_combine(params, partial: np.ndarray, state: np.ndarray) -> np.ndarray:
    partial = np.ascontiguousarray(partial, dtype=np.float32)
    rows = np.zeros((params.frames // max(1, params.row_frames)) * params.channels,
                    dtype=np.float32)
    _check(lib().oracle_ema_combine(C.byref(params), _a(partial), u64(partial.size),
                                    _a(state), _a(rows), u64(rows.size)), "ema_combine")
    return rows


def chain(desc: OChain, packed: np.ndarray, taps: np.ndarray, twiddles: np.ndarray,
          threads: int = 1, ring: np.ndarray | None = None,
          partial: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray | None]:
    """The plain chain over absolute blocks first_block .. + block_count.
    Returns (channel ring float32 pairs M x out_ring_blocks, EMA partials);
    `ring` / `partial` may be passed in preallocated."""
    taps = np.ascontiguousarray(taps, dtype=np.float32)
    tw = f32c(twiddles)
    M = desc.channels
    if ring is None:
        ring = np.zeros(2 * M * desc.out_ring_blocks, dtype=np.float32)
    np_ = 0
    if desc.frame_blocks:
        if partial is None:
            partial = np.zeros((desc.block_count // desc.frame_blocks) * M, dtype=np.float32)
        np_ = partial.size
    _check(lib().oracle_chain(C.byref(desc), _a(packed), u64(packed.nbytes), _a(taps),
                              u64(taps.size), _a(tw), u64(tw.size // 2), _a(ring),
                              u64(ring.size // 2),
                              _a(partial) if partial is not None else None, u64(np_),
                              i32(threads)), "chain")
    return ring, partial


def fp64_channelize(x: np.ndarray, h: np.ndarray, M: int, D: int, m0: int,
                    blocks: int) -> np.ndarray:
    """Direct FP64 channelizer definition; returns complex128 [blocks, M]."""
    xd = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
    hd = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(2 * blocks * M, dtype=np.float64)
    _check(lib().fp64_channelize(_a(xd), u64(xd.size // 2), _a(hd), u32(hd.size), u32(M),
                                 u32(D), u64(m0), u32(blocks), _a(out)), "fp64_channelize")
    return out.view(np.complex128).reshape(blocks, M)


def fp64_dft(x: np.ndarray) -> np.ndarray:
    xd = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
    out = np.zeros_like(xd)
    _check(lib().fp64_dft(_a(xd), u32(xd.size // 2), _a(out)), "fp64_dft")
    return out.view(np.complex128)


def vrx_fine(config, params, channel_ring: np.ndarray, taps: np.ndarray, nco: np.ndarray,
             out_samples: int) -> np.ndarray:
    """reference_vrx_fine; returns the fine ring (float32 pairs, out_samples)."""
    ch, tp, nc = f32c(channel_ring), f32c(taps), f32c(nco)
    out = np.zeros(2 * out_samples, dtype=np.float32)
    _check(lib().oracle_vrx_fine(C.byref(config), C.byref(params), _a(ch), u64(ch.size // 2),
                                 _a(tp), u64(tp.size // 2), _a(nc), u64(nc.size // 2), _a(out),
                                 u64(out_samples)), "vrx_fine")
    return out


def vrx_demod(config, params, fine_ring: np.ndarray, weights: np.ndarray) -> np.ndarray:
    fr = f32c(fine_ring)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.zeros(params.count * config.channels, dtype=np.float32)
    _check(lib().oracle_vrx_demod(C.byref(config), C.byref(params), _a(fr), u64(fr.size // 2),
                                  _a(w), u64(w.size), _a(out), u64(out.size)), "vrx_demod")
    return out


def vrx_carrier(config, params, nco: np.ndarray, fine_ring: np.ndarray,
                state: np.ndarray) -> None:
    """reference_vrx_carrier, in place on fine_ring (float32 pairs) and state
    (eight uint32 words)."""
    nc = f32c(nco)
    if fine_ring.dtype != np.float32 or not fine_ring.flags.c_contiguous:
        raise ValueError("fine ring must be contiguous float32 pairs")
    if state.dtype != np.uint32 or state.size != 8:
        raise ValueError("state must be eight uint32 words")
    _check(lib().oracle_vrx_carrier(C.byref(config), C.byref(params), _a(nc), u64(nc.size // 2),
                                    _a(fine_ring), u64(fine_ring.size // 2), _a(state),
                                    u64(state.size)), "vrx_carrier")


class OIqMomentsParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("src_offset", u32), ("count", u32)]


class OIqCorrectParams(C.Structure):
    _fields_ = [("ring_mask", u32), ("offset", u32), ("count", u32), ("dc_i", C.c_float),
                ("dc_q", C.c_float), ("cross", C.c_float), ("scale", C.c_float)]


def iq_moments(ring_mask: int, src_offset: int, count: int, ring: np.ndarray) -> np.ndarray:
    """reference_iq_moments: (chunks, 5) float32 chunk sums."""
    r = f32c(ring)
    chunks = (count + 63) // 64
    out = np.zeros(chunks * 5, dtype=np.float32)
    p = OIqMomentsParams(ring_mask, src_offset, count)
    _check(lib().oracle_iq_moments(C.byref(p), _a(r), u64(r.size // 2), _a(out), u64(out.size)),
           "iq_moments")
    return out.reshape(chunks, 5)


def iq_correct(ring_mask: int, offset: int, count: int, dc_i: float, dc_q: float, cross: float,
               scale: float, ring: np.ndarray) -> np.ndarray:
    """reference_iq_correct on a copy of the ring (float32 pairs)."""
    r = f32c(ring)
    out = r.copy()
    p = OIqCorrectParams(ring_mask, offset, count, dc_i, dc_q, cross, scale)
    _check(lib().oracle_iq_correct(C.byref(p), _a(r), u64(r.size // 2), _a(out),
                                   u64(out.size // 2)), "iq_correct")
    return out


def excise_segment(config, params, raw: np.ndarray, seg_power: np.ndarray) -> None:
    """reference_excise_segment, into seg_power (float32, seg ring) in place.
    The library's rvx_excision_check is the caller's to run first."""
    _check(lib().oracle_excise_segment(C.byref(config), C.byref(params), _a(f32c(raw)),
                                       _a(seg_power)), "excise_segment")


def excise_reference(config, params, seg_power: np.ndarray, seg_level: np.ndarray) -> None:
    _check(lib().oracle_excise_reference(C.byref(config), C.byref(params), _a(seg_power),
                                         _a(seg_level)), "excise_reference")


def excise_flag(config, params, raw: np.ndarray, seg_level: np.ndarray,
                flags: np.ndarray) -> None:
    _check(lib().oracle_excise_flag(C.byref(config), C.byref(params), _a(f32c(raw)),
                                    _a(seg_level), _a(flags)), "excise_flag")


def excise_apply(config, params, raw: np.ndarray, flags: np.ndarray, cleaned: np.ndarray,
                 removed: np.ndarray, counters: np.ndarray) -> None:
    """cleaned may be raw itself (in place, as the engine binds it)."""
    for a, dt in ((raw, np.float32), (cleaned, np.float32), (removed, np.float32),
                  (flags, np.uint32), (counters, np.uint32)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise ValueError("excise_apply wants contiguous float32 rings and uint32 words")
    _check(lib().oracle_excise_apply(C.byref(config), C.byref(params), _a(raw), _a(flags),
                                     _a(cleaned), _a(removed), _a(counters)), "excise_apply")


def comb_project(params, raw: np.ndarray, lines, table: np.ndarray) -> np.ndarray:
    """reference_comb_project: lines * sub_blocks complex64, line-major.
    lines: a ctypes array of the library's CombLine records."""
    r, t = f32c(raw), f32c(table)
    out = np.zeros(2 * params.lines * params.sub_blocks, dtype=np.float32)
    _check(lib().oracle_comb_project(C.byref(params), _a(r), C.cast(lines, vp), _a(t), _a(out)),
           "comb_project")
    return out.view(np.complex64)


def comb_subtract(params, raw: np.ndarray, lines, table: np.ndarray,
                  removed_samples: int) -> tuple[np.ndarray, np.ndarray]:
    """reference_comb_subtract on a copy of the ring: (cleaned, removed ring)."""
    r, t = f32c(raw), f32c(table)
    cleaned = r.copy()
    removed = np.zeros(2 * removed_samples, dtype=np.float32)
    _check(lib().oracle_comb_subtract(C.byref(params), _a(r), C.cast(lines, vp) if lines is not
                                      None else None, _a(t), _a(cleaned), _a(removed)),
           "comb_subtract")
    return cleaned, removed


def blank_detect(config, params, channel: np.ndarray, flags: np.ndarray) -> None:
    """reference_blank_detect into flags (float32, out_mask + 1) in place. The
    library's rvx_blanker_check is the caller's to run first."""
    if flags.dtype != np.float32 or not flags.flags.c_contiguous:
        raise ValueError("flags are contiguous float32")
    _check(lib().oracle_blank_detect(C.byref(config), C.byref(params), _a(f32c(channel)),
                                     _a(flags)), "blank_detect")


def blank_apply(config, params, channel: np.ndarray, flags: np.ndarray,
                blanked: np.ndarray) -> None:
    if blanked.dtype != np.float32 or not blanked.flags.c_contiguous:
        raise ValueError("blanked is contiguous float32 pairs")
    _check(lib().oracle_blank_apply(C.byref(config), C.byref(params), _a(f32c(channel)),
                                    _a(np.ascontiguousarray(flags, dtype=np.float32)),
                                    _a(blanked)), "blank_apply")


def noise_line(config, params, audio: np.ndarray, state: np.ndarray) -> None:
    """reference_line in place on audio (params.count floats) and state
    (float32, 24 + taps + history). The library's rvx_line_check is the
    caller's to run first."""
    for a in (audio, state):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise ValueError("audio and state are contiguous float32")
    if audio.size < params.count or state.size < 24 + config.taps + config.history:
        raise ValueError("audio or state shorter than the parameters")
    _check(lib().oracle_noise_line(C.byref(config), C.byref(params), _a(audio), _a(state)),
           "noise_line")


def noise_spectral(frame: int, params, tables: np.ndarray, audio: np.ndarray,
                   state: np.ndarray) -> None:
    """reference_spectral in place on audio and state (float32,
    2 + 3 frame + 2 (frame / 2 + 1))."""
    for a in (audio, state):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise ValueError("audio and state are contiguous float32")
    n = frame
    if audio.size < params.count or state.size < 2 + 2 * n + 2 * (n // 2 + 1) or \
            tables.size < 4 * n:
        raise ValueError("a buffer is shorter than the parameters")
    _check(lib().oracle_noise_spectral(C.c_uint32(frame), C.byref(params), _a(f32c(tables)),
                                       _a(audio), _a(state)), "noise_spectral")


def detect(config, average: np.ndarray):
    """The plain detector's decision on an averaged linear spectrum: returns
    (bands as (start, width, deflection) tuples, floor, sigma, peaks found)."""
    avg = np.ascontiguousarray(average, dtype=np.float64)
    fl = np.zeros(avg.size)
    sg = np.zeros(avg.size)
    cap = avg.size
    bands = (OBand * cap)()
    nb, npk = u32(), u32()
    _check(lib().oracle_detect(C.byref(config), _a(avg), _a(fl), _a(sg), bands, u32(cap),
                               C.byref(nb), C.byref(npk)), "detect")
    return [(b.start, b.width, b.deflection) for b in bands[:nb.value]], fl, sg, npk.value


def fp64_chain(codes: np.ndarray, taps: np.ndarray, M: int, D: int, first_block: int,
               blocks: int) -> np.ndarray:
    """The channelizer in float64 by branch sums, numpy's FFT and the output
    correction: complex128 [blocks, M], channel k of block m equal to
    sum_n h[n] x[mD - n] e^{-j 2 pi k (mD - n) / M} (pfb_branch.comp:9-19).
    codes: cs16 words (uint32) from sample 0."""
    c = np.ascontiguousarray(codes).view(np.int16).astype(np.float64) * 2.0 ** -15
    x = c[0::2] + 1j * c[1::2]
    L = taps.size // M
    h = taps.astype(np.float64).reshape(L, M)  # h[q][r] = taps[q*M + r]
    r = np.arange(M)
    slot = (M - r) % M
    k = np.arange(M)
    out = np.empty((blocks, M), dtype=np.complex128)
    for b in range(blocks):
        m = first_block + b
        idx = m * D - np.arange(L)[:, None] * M - r[None, :]
        u = (h * x[idx]).sum(axis=0)
        y = np.empty(M, dtype=np.complex128)
        y[slot] = u
        X = np.fft.fft(y)
        dm = (m * D) % M
        out[b] = X * np.exp(-2j * np.pi * ((k * dm) % M) / M)
    return out


def e4m3_decode(b: np.ndarray) -> np.ndarray:
    """OCP e4m3 bytes -> float64 (0x7F / 0xFF: NaN)."""
    b = np.asarray(b, dtype=np.uint8).astype(np.int64)
    s = np.where(b & 0x80, -1.0, 1.0)
    e = (b >> 3) & 15
    m = b & 7
    v = np.where(e == 0, m / 8.0 * 2.0 ** -6, (1 + m / 8.0) * 2.0 ** (e - 7.0))
    v = np.where((b & 0x7F) == 0x7F, np.nan, v)
    return s * v
