"""Device-side operators over torch CUDA tensors, through the C ABI.

``Device`` owns one ``rvx_ctx`` bound to torch's current stream, so the
library's launches are ordered with torch's own copies and allocations.
Complex sample spans are float32 tensors with a trailing dimension of 2 (or
complex64 viewed as such); every size the library checks comes from the
tensors themselves.

There is no fallback: without ``librvx_hip.so`` or without a GPU these raise.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .dsp import GridParams


def _p(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device operators take CUDA tensors")
    if not t.is_contiguous():
        raise ValueError("device operators take contiguous tensors")
    return t.data_ptr()


def _csamples(t: torch.Tensor) -> int:
    """Number of complex samples in a float32 (..., 2) / complex64 tensor."""
    if t.dtype == torch.complex64:
        return t.numel()
    if t.dtype != torch.float32:
        raise ValueError(f"complex span must be float32 pairs or complex64, got {t.dtype}")
    return t.numel() // 2


def _n(t: torch.Tensor) -> int:
    return t.numel()


class Device:
    """One context (core/gpu/context.h Context) on a HIP device."""

    def __init__(self, index: int = 0):
        self.lib = _lib.lib()
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU: the receive chain runs only on the device")
        self.index = index
        ctx = C.c_void_p()
        _lib.check(self.lib.rvx_ctx_create(index, C.byref(ctx)))
        self.ctx = ctx
        self.bind_stream()

    def bind_stream(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream(self.index)
        _lib.check(self.lib.rvx_ctx_set_stream(self.ctx, C.c_void_p(s.cuda_stream)))

    def close(self) -> None:
        if self.ctx:
            self.lib.rvx_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        _lib.check(self.lib.rvx_stream_sync(self.ctx))

    # -- operators ------------------------------------------------------
    def convert(self, fmt: int, packed: torch.Tensor, params: _lib.ConvertParams,
                ring: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_convert(self.ctx, fmt, _p(packed),
                                        packed.numel() * packed.element_size(),
                                        C.byref(params), _p(ring), _csamples(ring)))

    def cmul(self, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor) -> None:
        n = _csamples(a)
        if _csamples(b) != n or _csamples(out) != n:
            raise ValueError("cmul spans differ in length")
        _lib.check(self.lib.rvx_cmul(self.ctx, _p(a), _p(b), _p(out), n))

    def pfb_branch(self, grid: GridParams, taps: torch.Tensor, ring: torch.Tensor,
                   params: _lib.PfbBranchParams, out: torch.Tensor) -> None:
        g = grid.c()
        _lib.check(self.lib.rvx_pfb_branch(self.ctx, C.byref(g), _p(taps), _n(taps), _p(ring),
                                           _csamples(ring), C.byref(params), _p(out),
                                           _csamples(out)))

    def pfb_fft(self, params: _lib.PfbFftParams, branch: torch.Tensor, twiddles: torch.Tensor,
                ring: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_pfb_fft(self.ctx, C.byref(params), _p(branch),
                                        _csamples(branch), _p(twiddles), _csamples(twiddles),
                                        _p(ring), _csamples(ring)))

    def fft_radix2(self, twiddles: torch.Tensor, x: torch.Tensor, out: torch.Tensor, size: int,
                   batch: int) -> None:
        if _csamples(x) < size * batch or _csamples(out) < size * batch:
            raise ValueError("fft_radix2 spans shorter than size * batch")
        _lib.check(self.lib.rvx_fft_radix2(self.ctx, _p(twiddles), _csamples(twiddles), _p(x),
                                           _p(out), size, batch))

    def spectrum(self, params: _lib.SpectrumParams, ring: torch.Tensor, twiddles: torch.Tensor,
                 window: torch.Tensor, frame: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_spectrum(self.ctx, C.byref(params), _p(ring), _csamples(ring),
                                         _p(twiddles), _csamples(twiddles), _p(window),
                                         _n(window), _p(frame), _n(frame)))

    def spectrum_levels(self, params: _lib.SpectrumLevelsParams, frame: torch.Tensor,
                        levels: torch.Tensor) -> None:
        if _n(levels) < 2:
            raise ValueError("levels holds two floats")
        _lib.check(self.lib.rvx_spectrum_levels(self.ctx, C.byref(params), _p(frame),
                                                _n(frame), _p(levels)))

    def ema_partial(self, params: _lib.EmaParams, ring: torch.Tensor,
                    partial: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_ema_partial(self.ctx, C.byref(params), _p(ring),
                                            _csamples(ring), _p(partial), _n(partial)))

    def ema_combine(self, params: _lib.EmaParams, partial: torch.Tensor, state: torch.Tensor,
                    rows: torch.Tensor | None = None) -> None:
        if _n(state) < params.channels:
            raise ValueError("state holds one float per channel")
        _lib.check(self.lib.rvx_ema_combine(self.ctx, C.byref(params), _p(partial),
                                            _n(partial), _p(state), _p(rows),
                                            0 if rows is None else _n(rows)))

    def vrx_fine(self, config: _lib.VrxFineConfig, params: _lib.VrxFineParams,
                 channel_ring: torch.Tensor, taps: torch.Tensor, nco: torch.Tensor,
                 fine_ring: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_vrx_fine(self.ctx, C.byref(config), C.byref(params),
                                         _p(channel_ring), _csamples(channel_ring), _p(taps),
                                         _csamples(taps), _p(nco), _csamples(nco),
                                         _p(fine_ring), _csamples(fine_ring)))

    def vrx_demod(self, config: _lib.VrxDemodConfig, params: _lib.VrxDemodParams,
                  fine_ring: torch.Tensor, weights: torch.Tensor, audio: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_vrx_demod(self.ctx, C.byref(config), C.byref(params),
                                          _p(fine_ring), _csamples(fine_ring), _p(weights),
                                          _n(weights), _p(audio), _n(audio)))

    def iq_moments(self, params: _lib.IqMomentsParams, ring: torch.Tensor,
                   sums: torch.Tensor) -> None:
        """Five moments per 64-sample chunk of a ring block (iq_moments.comp)."""
        _lib.check(self.lib.rvx_iq_moments(self.ctx, C.byref(params), _p(ring), _csamples(ring),
                                           _p(sums), _n(sums)))

    def iq_moments_codes(self, fmt: int, codes: torch.Tensor, first: int, count: int,
                         sums: torch.Tensor) -> None:
        """The same sums over native codes, converted as rvx_convert would."""
        bps = 2 if fmt == 0 else 4
        n = codes.numel() * codes.element_size() // bps
        _lib.check(self.lib.rvx_iq_moments_codes(self.ctx, fmt, _p(codes), n, first, count,
                                                 _p(sums), _n(sums)))

    def iq_correct(self, params: _lib.IqCorrectParams, ring_in: torch.Tensor,
                   ring_out: torch.Tensor | None = None) -> None:
        """DC and I/Q correction of a ring block (iq_correct.comp); in place
        when ring_out is None."""
        out = ring_in if ring_out is None else ring_out
        _lib.check(self.lib.rvx_iq_correct(self.ctx, C.byref(params), _p(ring_in),
                                           _csamples(ring_in), _p(out), _csamples(out)))

    # -- per-receiver impulse blanker (noise_blank.comp) -------------------
    def blank_detect(self, config, params, channel_ring: torch.Tensor,
                     flags: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_blank_detect(self.ctx, C.byref(config), C.byref(params),
                                             _p(channel_ring), _csamples(channel_ring), _p(flags),
                                             _n(flags)))

    def blank_apply(self, config, params, channel_ring: torch.Tensor, flags: torch.Tensor,
                    blanked: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_blank_apply(self.ctx, C.byref(config), C.byref(params),
                                            _p(channel_ring), _csamples(channel_ring), _p(flags),
                                            _n(flags), _p(blanked), _csamples(blanked)))

    # -- line kernel and spectral noise reduction (noise_line/spectral.comp) --
    def noise_line(self, config, params, audio: torch.Tensor, state: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_noise_line(self.ctx, C.byref(config), C.byref(params), _p(audio),
                                           _n(audio), _p(state), _n(state)))

    def noise_line_batch(self, config, items: torch.Tensor, n_items: int) -> None:
        """items: a device byte tensor of n_items rvx_noise_line_item records,
        each checked with rvx_noise_line_item_check by the caller."""
        _lib.check(self.lib.rvx_noise_line_batch(self.ctx, C.byref(config), _p(items), n_items))

    def noise_spectral(self, frame: int, params, tables: torch.Tensor, audio: torch.Tensor,
                       state: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_noise_spectral(self.ctx, frame, C.byref(params), _p(tables),
                                               _n(tables), _p(audio), _n(audio), _p(state),
                                               _n(state)))

    def noise_spectral_batch(self, frame: int, tables: torch.Tensor, items: torch.Tensor,
                             n_items: int) -> None:
        _lib.check(self.lib.rvx_noise_spectral_batch(self.ctx, frame, _p(tables), _n(tables),
                                                     _p(items), n_items))

    # -- comb cancellation (comb_project.comp, comb_subtract.comp) --------
    def comb_project(self, params, raw: torch.Tensor, lines: torch.Tensor, n_lines: int,
                     table: torch.Tensor, projections: torch.Tensor) -> None:
        """lines: a device copy of rvx_comb_line records."""
        _lib.check(self.lib.rvx_comb_project(self.ctx, C.byref(params), _p(raw), _csamples(raw),
                                             _p(lines), n_lines, _p(table), _csamples(table),
                                             _p(projections), _csamples(projections)))

    def comb_subtract(self, params, raw: torch.Tensor, lines: torch.Tensor, n_lines: int,
                      table: torch.Tensor, cleaned: torch.Tensor, removed: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_comb_subtract(self.ctx, C.byref(params), _p(raw), _csamples(raw),
                                              _p(lines), n_lines, _p(table), _csamples(table),
                                              _p(cleaned), _csamples(cleaned), _p(removed),
                                              _csamples(removed)))

    # -- impulse excision (impulse_excise.comp's four passes) -------------
    def excise_segment(self, config, params, raw: torch.Tensor, seg_power: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_excise_segment(self.ctx, C.byref(config), C.byref(params),
                                               _p(raw), _csamples(raw), _p(seg_power),
                                               _n(seg_power)))

    def excise_reference(self, config, params, seg_power: torch.Tensor,
                         seg_level: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_excise_reference(self.ctx, C.byref(config), C.byref(params),
                                                 _p(seg_power), _p(seg_level),
                                                 min(_n(seg_power), _n(seg_level))))

    def excise_flag(self, config, params, raw: torch.Tensor, seg_level: torch.Tensor,
                    flags: torch.Tensor) -> None:
        _lib.check(self.lib.rvx_excise_flag(self.ctx, C.byref(config), C.byref(params), _p(raw),
                                            _csamples(raw), _p(seg_level), _n(seg_level),
                                            _p(flags), _n(flags)))

    def excise_apply(self, config, params, raw: torch.Tensor, flags: torch.Tensor,
                     cleaned: torch.Tensor, removed: torch.Tensor,
                     counters: torch.Tensor) -> None:
        """cleaned may be raw itself (in place)."""
        if counters.numel() < 4 or counters.element_size() != 4:
            raise ValueError("counters are four 32-bit words")
        _lib.check(self.lib.rvx_excise_apply(self.ctx, C.byref(config), C.byref(params), _p(raw),
                                             _csamples(raw), _p(flags), _n(flags), _p(cleaned),
                                             _csamples(cleaned), _p(removed), _csamples(removed),
                                             _p(counters)))

    def vrx_carrier(self, config: _lib.VrxCarrierConfig, params: _lib.VrxCarrierParams,
                    nco: torch.Tensor, fine_ring: torch.Tensor, state: torch.Tensor) -> None:
        """The carrier loop over one receiver's fine ring, in place; state is
        eight int32/uint32 words on the device."""
        if state.numel() != 8 or state.element_size() != 4:
            raise ValueError("carrier state is eight 32-bit words")
        _lib.check(self.lib.rvx_vrx_carrier(self.ctx, C.byref(config), C.byref(params), _p(nco),
                                            _csamples(nco), _p(fine_ring), _csamples(fine_ring),
                                            _p(state)))

    def vrx_carrier_batch(self, items: torch.Tensor, n_items: int, nco: torch.Tensor) -> None:
        """Many receivers' carrier loops on one NCO table; items is a device
        copy of rvx_vrx_carrier_item records (each checked on the host)."""
        if items.numel() * items.element_size() < n_items * C.sizeof(_lib.VrxCarrierItem):
            raise ValueError("item buffer smaller than n_items records")
        _lib.check(self.lib.rvx_vrx_carrier_batch(self.ctx, _p(items), n_items, _p(nco),
                                                  _csamples(nco)))

    def chain(self, desc: _lib.ChainDesc, packed: torch.Tensor, taps: torch.Tensor,
              twiddles: torch.Tensor, ring: torch.Tensor,
              partial: torch.Tensor | None = None) -> None:
        bps = {0: 2, 2: 4, 4: 8}.get(desc.format, 4)  # cu8, cs16, cf32
        if packed.numel() * packed.element_size() < desc.input_samples * bps:
            raise ValueError("input tensor smaller than input_samples")
        _lib.check(self.lib.rvx_chain(self.ctx, C.byref(desc), _p(packed), _p(taps), _n(taps),
                                      _p(twiddles), _csamples(twiddles), _p(ring),
                                      _csamples(ring), _p(partial),
                                      0 if partial is None else _n(partial)))

    def chain_mfma(self, desc: _lib.ChainDesc, precision: int, packed: torch.Tensor,
                   taps: torch.Tensor, twiddles: torch.Tensor, ring: torch.Tensor,
                   partial: torch.Tensor | None = None) -> None:
        """The chain with its transform on the matrix cores (not bit-exact)."""
        if packed.numel() * packed.element_size() < desc.input_samples * 4:
            raise ValueError("input tensor smaller than input_samples")
        _lib.check(self.lib.rvx_chain_mfma(self.ctx, C.byref(desc), precision, _p(packed),
                                           _p(taps), _n(taps), _p(twiddles),
                                           _csamples(twiddles), _p(ring), _csamples(ring),
                                           _p(partial), 0 if partial is None else _n(partial)))

    def spectrum_fp8(self, power: torch.Tensor, out: torch.Tensor) -> None:
        """e4m3 of 10 log10(power) / 4 (display spectrum / waterfall rows)."""
        if out.numel() < power.numel() or out.element_size() != 1:
            raise ValueError("out must hold one byte per power value")
        _lib.check(self.lib.rvx_spectrum_fp8(self.ctx, _p(power), power.numel(), _p(out)))

    # -- graphs ---------------------------------------------------------
    def graph_begin(self) -> None:
        """Start capturing this context's stream (not the null stream)."""
        _lib.check(self.lib.rvx_graph_begin(self.ctx))

    def graph_end(self) -> "Graph":
        g = C.c_void_p()
        _lib.check(self.lib.rvx_graph_end(self.ctx, C.byref(g)))
        return Graph(self, g)

    # -- timing ---------------------------------------------------------
    def timer_begin(self) -> None:
        _lib.check(self.lib.rvx_timer_begin(self.ctx))

    def timer_end(self) -> float:
        ms = C.c_float()
        _lib.check(self.lib.rvx_timer_end(self.ctx, C.byref(ms)))
        return ms.value


class Graph:
    """A captured, instantiated sequence of launches (HIP graph)."""

    def __init__(self, dev: Device, handle: C.c_void_p):
        self.dev = dev
        self.handle = handle

    def launch(self) -> None:
        _lib.check(self.dev.lib.rvx_graph_launch(self.dev.ctx, self.handle))

    def close(self) -> None:
        if self.handle:
            self.dev.lib.rvx_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
