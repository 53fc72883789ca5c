"""The streaming engine (rvx_engine_*): pushes of native samples in, channel
ring, averaged spectrum and per-receiver audio out.

Graph::on_block (core/engine/graph.cpp:4770-5480) and the receivers' stream
cursors (core/engine/vrx_stage.cpp:1119-1480), run natively in the library;
this is the Python face of it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .dsp import GridParams, VrxPlan
from .ops import Device


CALIBRATE_DC, CALIBRATE_IQ = 1, 2  # RVX_CALIBRATE_*


class Engine:
    def __init__(self, dev: Device, grid: GridParams, prototype: np.ndarray, fmt: int,
                 max_push: int, ring_blocks: int, frame_blocks: int = 0, row_frames: int = 1,
                 alpha: float = 1.0 / 16.0, start_index: int = 0, calibration: int = 0,
                 sample_rate: int = 0, ring_pitch: int = 0):
        """calibration: CALIBRATE_DC | CALIBRATE_IQ turns on the front-end
        correction stage (sample_rate in Hz then sets its time constants).
        ring_pitch: the channel ring's row stride in samples (0: ring_blocks,
        the reference's layout; ring_blocks + _lib.RING_PITCH_PAD spreads a
        block's writes over the HBM channels)."""
        self.dev = dev
        self.lib = _lib.lib()
        self.grid = grid
        self.fmt = fmt
        cfg = _lib.EngineConfig(fmt, grid.channels, grid.taps_per_branch, grid.decimation,
                                max_push, ring_blocks, frame_blocks, row_frames,
                                np.float32(alpha), calibration, start_index, sample_rate,
                                ring_pitch)
        self.ring_pitch = ring_pitch or ring_blocks
        taps = np.ascontiguousarray(prototype, dtype=np.float32)
        h = C.c_void_p()
        _lib.check(self.lib.rvx_engine_create(dev.ctx, C.byref(cfg), taps.ctypes.data_as(
            _lib.fptr), taps.size, C.byref(h)))
        self.handle = h
        self.channels = {}

    def add(self, plan: VrxPlan, audio_capacity: int, tile: int = 0, grid: int = 0) -> int:
        """A receiver on grid `grid` (0: the engine's own; else an id from
        add_grid, with the plan made on that grid)."""
        taps = np.ascontiguousarray(plan.fine_taps.view(np.float32))
        w = np.ascontiguousarray(plan.weights, dtype=np.float32)
        rid = _lib.u32()
        _lib.check(self.lib.rvx_engine_add_vrx_on(self.handle, grid, C.byref(plan.plan),
                                                  taps.ctypes.data_as(_lib.fptr), taps.size,
                                                  w.ctypes.data_as(_lib.fptr), w.size,
                                                  audio_capacity, tile, C.byref(rid)))
        self.channels[rid.value] = plan.demod.channels
        return rid.value

    def add_grid(self, grid: GridParams, prototype: np.ndarray, ring_blocks: int,
                 max_rows: int, ring_pitch: int = 0) -> int:
        """A second grid on the same stream (rvx_engine_add_grid) storing up
        to max_rows of its channels, rows ring_pitch samples apart (0:
        ring_blocks); returns its id."""
        cfg = _lib.EngineGrid(grid.channels, grid.taps_per_branch, grid.decimation, ring_blocks,
                              max_rows, ring_pitch)
        taps = np.ascontiguousarray(prototype, dtype=np.float32)
        gid = _lib.u32()
        _lib.check(self.lib.rvx_engine_add_grid(self.handle, C.byref(cfg),
                                                taps.ctypes.data_as(_lib.fptr), taps.size,
                                                C.byref(gid)))
        return gid.value

    def set_excision(self, config: _lib.ExcisionConfig, threshold: float) -> None:
        """Impulse excision before the channelizer (rvx_engine_set_excision);
        before the first push."""
        _lib.check(self.lib.rvx_engine_set_excision(self.handle, C.byref(config), threshold))

    def set_comb(self, lines, phase_index: int = 0, subtract: bool = True) -> None:
        """Comb cancellation (rvx_engine_set_comb): lines is a ctypes array
        of CombLine (phases at absolute sample phase_index) or empty."""
        n = len(lines) if lines is not None else 0
        ptr = C.cast(lines, C.POINTER(_lib.CombLine)) if n else None
        _lib.check(self.lib.rvx_engine_set_comb(self.handle, ptr, n, phase_index,
                                                1 if subtract else 0))

    def comb_projections(self):
        """(projections complex64 [lines, sub_blocks], first sample, samples)
        of the last push."""
        out = np.zeros(2 * 1024 * 16, dtype=np.float32)
        first, count, lines, subs = _lib.u64(), _lib.u64(), _lib.u32(), _lib.u32()
        _lib.check(self.lib.rvx_engine_comb_projections(
            self.handle, out.ctypes.data_as(_lib.fptr), out.size, C.byref(first), C.byref(count),
            C.byref(lines), C.byref(subs)))
        n = lines.value * subs.value
        proj = out[:2 * n].view(np.complex64).reshape(lines.value, subs.value).copy()
        return proj, first.value, count.value

    def excision_counters(self) -> np.ndarray:
        out = np.zeros(4, dtype=np.uint32)
        _lib.check(self.lib.rvx_engine_excision_counters(
            self.handle, out.ctypes.data_as(C.POINTER(_lib.u32))))
        return out

    def grid_ring(self, grid: int) -> int:
        return self.lib.rvx_engine_grid_ring(self.handle, grid)

    def grid_row(self, grid: int, channel: int) -> tuple[int, int]:
        """(ring row holding `channel` on `grid`, first block written into it)."""
        row, first = _lib.u32(), _lib.u64()
        _lib.check(self.lib.rvx_engine_grid_row(self.handle, grid, channel, C.byref(row),
                                                C.byref(first)))
        return row.value, first.value

    def remove(self, rid: int) -> None:
        _lib.check(self.lib.rvx_engine_remove_vrx(self.handle, rid))
        self.channels.pop(rid, None)

    def push(self, codes) -> None:
        """Native codes: a numpy array (host) or a CUDA tensor (device)."""
        bps = 4 if self.fmt == 2 else 2
        if isinstance(codes, np.ndarray):
            a = np.ascontiguousarray(codes)
            _lib.check(self.lib.rvx_engine_push(self.handle, a.ctypes.data, a.nbytes // bps, 0))
        else:
            if not codes.is_cuda or not codes.is_contiguous():
                raise ValueError("device pushes take contiguous CUDA tensors")
            n = codes.numel() * codes.element_size() // bps
            _lib.check(self.lib.rvx_engine_push(self.handle, codes.data_ptr(), n, 1))

    def read_audio(self, rid: int, max_values: int = 1 << 24) -> tuple[int, np.ndarray]:
        out = np.empty(max_values, dtype=np.float32)
        first, n = _lib.u64(), _lib.u64()
        _lib.check(self.lib.rvx_engine_read_audio(self.handle, rid, out.ctypes.data_as(_lib.fptr),
                                                  max_values, C.byref(first), C.byref(n)))
        return first.value, out[:n.value].copy()

    def spectrum(self) -> np.ndarray:
        out = np.empty(self.grid.channels, dtype=np.float32)
        _lib.check(self.lib.rvx_engine_spectrum(self.handle, out.ctypes.data_as(_lib.fptr),
                                                out.size))
        return out

    def channel_ring(self) -> int:
        return self.lib.rvx_engine_channel_ring(self.handle)

    def stats(self) -> _lib.EngineStats:
        s = _lib.EngineStats()
        _lib.check(self.lib.rvx_engine_stats_get(self.handle, C.byref(s)))
        return s

    def close(self) -> None:
        if self.handle:
            self.lib.rvx_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_display(self, rid: int, display_ring: int = 1 << 14, transform: int = 1024) -> None:
        """The receiver's display tap (rvx_engine_set_display); display_ring
        0 turns it off."""
        _lib.check(self.lib.rvx_engine_set_display(self.handle, rid, display_ring, transform))

    def passband(self, rid: int, transform: int = 1024):
        """The pane over the newest `transform` display samples:
        (dB values ascending in frequency, one past the window's last display
        output, the display rate); None before there are enough samples."""
        out = np.empty(transform // 2, dtype=np.float32)
        end, rate = _lib.u64(), _lib.i64()
        rc = self.lib.rvx_engine_passband(self.handle, rid, out.ctypes.data, out.size,
                                          C.byref(end), C.byref(rate))
        if rc == -4:
            return None
        _lib.check(rc)
        return out, end.value, rate.value

    def set_blanker(self, rid: int, config: _lib.BlankerConfig | None,
                    threshold: float = 0.0) -> None:
        """The receiver's impulse blanker on its coarse channel
        (rvx_engine_set_blanker); None turns it off."""
        _lib.check(self.lib.rvx_engine_set_blanker(
            self.handle, rid, C.byref(config) if config is not None else None, threshold))

    def set_noise(self, rid: int, line_config: _lib.LineConfig | None = None,
                  line_params: _lib.LineParams | None = None, spectral_frame: int = 0,
                  spectral_params: _lib.SpectralParams | None = None) -> None:
        """A mono receiver's audio noise stages (rvx_engine_set_noise): the
        line kernel and / or spectral reduction, run on every push's audio;
        None turns a stage off. Restarts both stages' state."""
        if line_params is not None and line_config is None:
            line_config = _lib.LineConfig()
        _lib.check(self.lib.rvx_engine_set_noise(
            self.handle, rid, C.byref(line_config) if line_config is not None else None,
            C.byref(line_params) if line_params is not None else None, spectral_frame,
            C.byref(spectral_params) if spectral_params is not None else None))

    def set_noise_plan(self, rid: int, plan, line_config: _lib.LineConfig | None = None,
                       spectral_frame: int = 0, blanker: _lib.BlankerConfig | None = None) -> None:
        """From revenant_amd.noise.plan_noise's NoisePlan; the blanker needs
        its design for the receiver's channel rate (dsp.design_blanker)."""
        if plan.blank:
            if blanker is None:
                raise ValueError("the plan blanks: pass design_blanker(channel_rate)")
            self.set_blanker(rid, blanker, plan.blank_threshold)
        else:
            self.set_blanker(rid, None)
        self.set_noise(rid, line_config if plan.line else None,
                       plan.line_params if plan.line else None,
                       spectral_frame if plan.spectral else 0,
                       plan.spectral_params if plan.spectral else None)

    def front_end(self) -> tuple[_lib.FrontEndEstimate, _lib.IqCorrectParams]:
        """The front-end estimate and the correction the newest push took."""
        est, applied = _lib.FrontEndEstimate(), _lib.IqCorrectParams()
        _lib.check(self.lib.rvx_engine_front_end(self.handle, C.byref(est), C.byref(applied)))
        return est, applied
