"""Batched receivers: descriptors, the work list, tile tuning and the launch.

The reference records two or three dispatches per receiver and block
(`DemodStage::record`, core/engine/vrx_stage.cpp:1306-1399): vrx_fine into a
fine ring, then vrx_demod from it. Here every receiver of a step runs in one
`rvx_vrx_batch` launch. A persistent grid takes (receiver, audio tile) items
from a queue; each item computes its fine samples, detector values and audio
on chip, so the fine ring never exists in HBM. The result is bit-identical to
`rvx_vrx_fine` followed by `rvx_vrx_demod` for the same plan and indices.

A receiver's block of audio outputs a0 .. a0+count-1 needs fine outputs
F0 .. F0 + (count-1)*R + hist with F0 = a0*R - hist
(`demod_fine_history`, core/dsp/vrx_reference.h:555); the descriptor carries
the fine-block params of exactly that range (`resample_block`,
vrx_reference.cpp:2786-2871) and the demod-block params of the audio
(vrx_reference.cpp:2873-2909).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .dsp import VrxPlan
from .ops import Device, _csamples, _p

# The batching workload's receiver mix: (mode name, count, bandwidth Hz).
STANDARD_MIX = (("nfm", 64, 12500), ("am", 64, 10000), ("wfm", 32, 200_000),
                ("usb", 64, 2700), ("cw", 32, 500))


def standard_mix(rate: int, narrow_grid, wide_grid, seed: int, audio_rate: int = 48000,
                 mix=STANDARD_MIX):
    """The mix's receivers at seeded uniform offsets within +-0.45 rate, placed
    and planned: 200 kHz receivers on ``wide_grid`` (the grid channel_count_for
    picks for them), the rest on ``narrow_grid``. Returns ([(name, plan)], the
    sorted wide-grid channels the wide receivers use)."""
    from .dsp import (DEMOD_AM, DEMOD_CW, DEMOD_NFM, DEMOD_USB, DEMOD_WFM, VrxRequest, place,
                      plan_vrx)

    modes = {"nfm": DEMOD_NFM, "am": DEMOD_AM, "wfm": DEMOD_WFM, "usb": DEMOD_USB,
             "cw": DEMOD_CW}
    g = np.random.default_rng(seed)
    plans = []
    for name, count, bw in mix:
        for _ in range(count):
            f = int(g.integers(-int(0.45 * rate), int(0.45 * rate)))
            req = VrxRequest(f, modes[name], bw, audio_rate, cw_pitch=700 if name == "cw" else 0)
            grid = wide_grid if name == "wfm" else narrow_grid
            plans.append((name, plan_vrx(grid, rate, req, place(grid, rate, req))))
    wide_set = sorted({p.plan.placement.channel for n, p in plans if n == "wfm"})
    return plans, wide_set

MAX_SPAN = _lib.VRX_TILE_SPAN  # fine samples one tile holds on chip
_UNBOUNDED = 0xFFFFFFFF         # ring masks of the on-chip fine block: no wrap


def max_tile(plan: VrxPlan) -> int:
    """Largest audio tile whose fine span fits the on-chip budget."""
    r, hist = plan.demod.decimation, plan.history
    if hist + 1 > MAX_SPAN:
        raise ValueError(f"a receiver whose filters reach {hist} fine samples does not fit a tile")
    return (MAX_SPAN - hist - 1) // r + 1


@dataclass
class Receiver:
    """One receiver of the batch: its plan and where its channel sits."""

    plan: VrxPlan
    channel_base: int    # cf32 index of the channel's first slot in the channel ring
    channel_mask: int    # channel length - 1 (a power of two minus one)
    ring: torch.Tensor   # the channel ring it reads (cf32 as float32 pairs)
    taps: torch.Tensor = field(default=None)      # device copies of the plan's tables
    weights: torch.Tensor = field(default=None)

    def upload(self, device: torch.device) -> None:
        if self.taps is None:
            self.taps = torch.from_numpy(self.plan.fine_taps.view(np.float32).copy()).to(device)
        if self.weights is None:
            self.weights = torch.from_numpy(self.plan.weights.copy()).to(device)

    def audio_range(self, first_sample: int, last_sample: int) -> tuple[int, int]:
        """(first_audio, count): the audio outputs whose every input lies in
        channel samples first_sample .. last_sample (absolute indices).
        Fine output j reads samples floor(j*Fc/Fd) - (taps-1) .. floor(j*Fc/Fd)
        (vrx_reference.cpp:489-620); audio a reads fine a*R - hist .. a*R."""
        plan = self.plan
        fc, fd = plan.plan.channel_rate, plan.plan.demod_rate
        r, hist, taps = plan.demod.decimation, plan.history, plan.fine.taps
        # lowest fine j with floor(j*Fc/Fd) >= first_sample + taps - 1
        lo_n = first_sample + taps - 1
        j_lo = -(-lo_n * fd // fc)
        # highest fine j with floor(j*Fc/Fd) <= last_sample
        j_hi = ((last_sample + 1) * fd - 1) // fc
        a_lo = -(-(j_lo + hist) // r)
        a_hi = j_hi // r
        return a_lo, max(0, a_hi - a_lo + 1)

    def first_input(self, first_audio: int) -> int:
        """Lowest channel sample index the audio block from first_audio reads."""
        f0 = first_audio * self.plan.demod.decimation - self.plan.history
        _, n0 = self.plan.fine_block(self.channel_base, self.channel_mask, _UNBOUNDED, f0, 1)
        return n0 - (self.plan.fine.taps - 1)


class ReceiverBatch:
    """Every receiver of a step in one launch.

    `prepare` records the step (rvx_vrx_batch_plan: descriptors split where
    the 32-bit resampler recurrence requires, work ordered largest first) and
    writes it to the device; `launch` is the one kernel launch (plus the queue
    reset), safe to capture in a HIP graph.
    """

    def __init__(self, dev: Device, receivers: list[Receiver], nco: torch.Tensor,
                 workgroups: int = 0):
        self.dev = dev
        self.receivers = receivers
        self.nco = nco
        self.workgroups = workgroups
        device = nco.device
        for r in receivers:
            r.upload(device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=device)
        self.tiles = [max_tile(r.plan) for r in receivers]
        self._descs = None
        self._work = None
        self.n_desc = 0
        self.n_work = 0

    def prepare(self, first_audio: list[int], counts: list[int], audio: list[torch.Tensor],
                tiles: list[int] | None = None) -> None:
        if tiles is not None:
            self.tiles = [min(t, max_tile(r.plan)) for t, r in zip(tiles, self.receivers)]
        spans = (_lib.VrxSpan * len(self.receivers))()
        for i, (r, a0, n, out, t) in enumerate(zip(self.receivers, first_audio, counts, audio,
                                                  self.tiles)):
            if out.numel() < n * r.plan.demod.channels:
                raise ValueError("audio tensor smaller than count * channels")
            spans[i] = _lib.VrxSpan(C.pointer(r.plan.plan), r.channel_base, r.channel_mask,
                                    _p(r.ring), _p(r.taps), _p(r.weights), _p(out), a0, n, t)
        lib = _lib.lib()
        nd, nw = _lib.u32(), _lib.u32()
        _lib.check(lib.rvx_vrx_batch_plan(spans, len(spans), None, 0, C.byref(nd), None, 0,
                                          C.byref(nw)))
        if nd.value == 0:
            raise ValueError("no receiver has audio to compute")
        descs = (_lib.VrxDesc * nd.value)()
        work = (_lib.VrxWork * nw.value)()
        _lib.check(lib.rvx_vrx_batch_plan(spans, len(spans), descs, nd.value, C.byref(nd), work,
                                          nw.value, C.byref(nw)))
        device = self.nco.device
        self.host_descs = descs
        self._descs = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(device)
        self._work = torch.frombuffer(bytearray(bytes(work)), dtype=torch.uint8).to(device)
        self.n_desc = nd.value
        self.n_work = nw.value

    def launch(self) -> None:
        if self._descs is None:
            raise RuntimeError("prepare() the batch first")
        _lib.check(self.dev.lib.rvx_vrx_batch(
            self.dev.ctx, _p(self._descs), self.n_desc, _p(self._work), self.n_work,
            _p(self.counter), _p(self.nco), self.nco.numel() // 2, self.workgroups))

    def unbatched(self, fine_ring: torch.Tensor) -> None:
        """The same work as launch(), recorded the way the reference records it
        (core/engine/vrx_stage.cpp:1306-1399): per receiver segment, rvx_vrx_fine
        into a fine ring in HBM, then rvx_vrx_demod from it. Same audio bits;
        two launches per segment instead of one for everything."""
        lib = self.dev.lib
        cap = _csamples(fine_ring)
        if cap & (cap - 1):
            raise ValueError("the fine ring holds a power of two of samples")
        owner = {r.taps.data_ptr(): r for r in self.receivers}
        for d in self.host_descs:
            r = owner[d.taps]
            fp = _lib.VrxFineParams.from_buffer_copy(d.fine)
            if fp.count > cap:
                raise ValueError(f"a segment of {fp.count} fine samples exceeds the ring")
            fp.out_offset, fp.out_mask = 0, cap - 1
            dp = _lib.VrxDemodParams.from_buffer_copy(d.demod)
            dp.in_offset, dp.in_mask = d.hist, cap - 1
            _lib.check(lib.rvx_vrx_fine(self.dev.ctx, C.byref(d.fine_config), C.byref(fp),
                                        d.channel_ring, _csamples(r.ring), d.taps,
                                        _csamples(r.taps), _p(self.nco), self.nco.numel() // 2,
                                        _p(fine_ring), cap))
            _lib.check(lib.rvx_vrx_demod(self.dev.ctx, C.byref(d.demod_config), C.byref(dp),
                                         _p(fine_ring), cap, d.weights, r.weights.numel(),
                                         d.audio, d.demod.count * d.demod_config.channels))

    def tune(self, first_audio: list[int], counts: list[int], audio: list[torch.Tensor],
             candidates=(32, 64, 128, 256, 512, 1024, 2048), reps: int = 3) -> list[int]:
        """Per mode class, the audio tile that minimises the whole batch's time.

        Coordinate descent over receiver classes (mode, decimation, filter
        lengths): each class's tile is swept with the others held, keeping the
        fastest. Tiles change only the work split, never the output bits.
        """
        classes: dict[tuple, list[int]] = {}
        for v, r in enumerate(self.receivers):
            d, f = r.plan.demod, r.plan.fine
            classes.setdefault((d.mode, d.decimation, d.audio_taps, d.dc_taps, f.taps), []).append(v)
        tiles = list(self.tiles)

        def timed(ts: list[int]) -> float:
            self.prepare(first_audio, counts, audio, ts)
            self.launch()
            best = float("inf")
            for _ in range(reps):
                self.dev.timer_begin()
                self.launch()
                best = min(best, self.dev.timer_end())
            return best

        best_ms = timed(tiles)
        for members in classes.values():
            cap = max_tile(self.receivers[members[0]].plan)
            for t in candidates:
                if t > cap or t == tiles[members[0]]:
                    continue
                trial = list(tiles)
                for v in members:
                    trial[v] = t
                ms = timed(trial)
                if ms < best_ms:
                    best_ms, tiles = ms, trial
        self.prepare(first_audio, counts, audio, tiles)
        return tiles
