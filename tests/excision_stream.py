"""A capture through impulse excision block by block, the way the engine
drives the four passes (tests/reference/test_impulse_excision.cpp:113-196):
the cursor plans each block, the passes run in order, the apply pass cleans
the ring in place, and each block's removed values are copied out of the
short removed ring before the next block reuses it. `passes` is the oracle or
the device.
"""
from __future__ import annotations

import math

import numpy as np

from revenant_amd import _lib
from revenant_amd.excision import (APPLY, FLAG, REFERENCE, SEGMENT, ExcisionCursor,
                                   design_excision, excision_threshold, size_excision_rings)


def impulsive_capture(length: int, seed: int, tone: float, cycles: float, impulses: int,
                      max_width: int, where: list | None = None) -> np.ndarray:
    """Complex Gaussian noise (0.01 per component), a tone, and impulses of
    heights 1..3 and widths 1..max_width, (h, -0.7h) added."""
    g = np.random.default_rng(seed)
    n = np.arange(length, dtype=np.float64)
    z = (g.normal(0.0, 0.01, length) + 1j * g.normal(0.0, 0.01, length)) + \
        tone * np.exp(2j * math.pi * cycles * n)
    out = z.astype(np.complex64)
    for k in range(impulses):
        at = int(g.integers(length // 8, length - 1024))
        width = 1 + k % max_width
        h = 1.0 + 0.5 * (k % 5)
        out[at:at + width] += np.complex64(h - 0.7j * h)
        if where is not None:
            where.append(at)
    return out


class OraclePasses:
    def __init__(self):
        from oracle import oracle
        self.o = oracle

    def alloc(self, ring, seg, flags, removed):
        return (np.ascontiguousarray(ring.view(np.float32).copy()),
                np.zeros(seg, np.float32), np.zeros(seg, np.float32), np.zeros(flags, np.uint32),
                np.zeros(2 * removed, np.float32), np.zeros(4, np.uint32))

    def run(self, cfg, plan, st):
        raw, power, level, flags, removed, counters = st
        c = _lib.ExcisionConfig.from_buffer_copy(bytes(cfg))
        if plan.segment.count:
            c.pass_ = SEGMENT
            self.o.excise_segment(c, plan.segment, raw, power)
        if plan.reference.count:
            c.pass_ = REFERENCE
            self.o.excise_reference(c, plan.reference, power, level)
        if plan.flag.count:
            c.pass_ = FLAG
            self.o.excise_flag(c, plan.flag, raw, level, flags)
        if plan.apply.count:
            c.pass_ = APPLY
            counters[:] = 0
            self.o.excise_apply(c, plan.apply, raw, flags, raw, removed, counters)
        return counters.astype(np.uint64)

    def removed(self, st):
        return st[4]

    def ring(self, st):
        return st[0]


def run_stream(capture: np.ndarray, rate: int, block: int, threshold_db: float, passes,
               guard: int | None = None):
    cfg = design_excision(rate)
    if guard is not None:  # hang and lead overridden (the kernels take up to 16)
        cfg.hang = cfg.lead = guard
    ring_n = 1 << (capture.size - 1).bit_length()
    mask = ring_n - 1
    rings = size_excision_rings(cfg, block, 1, mask)
    full = np.zeros(ring_n, np.complex64)
    full[:capture.size] = capture
    st = passes.alloc(full, rings.seg_mask + 1, rings.flag_mask + 1, rings.removed_mask + 1)
    cursor = ExcisionCursor(cfg, mask, rings, excision_threshold(threshold_db))
    removed = np.full(capture.size, np.complex64(complex(-0.0, -0.0)))
    removed.view(np.float32)[:] = -0.0
    counters = np.zeros(4, np.uint64)
    cleaned_end = 0
    for start in range(0, capture.size, block):
        count = min(block, capture.size - start)
        plan = cursor.plan(start, count)
        counters += passes.run(cfg, plan, st)
        if plan.apply.count:
            rem = passes.removed(st).view(np.complex64)
            lo = plan.cleaned_end - plan.apply.count
            idx = np.arange(lo, plan.cleaned_end)
            removed[idx] = rem[idx & rings.removed_mask]
        cleaned_end = plan.cleaned_end
    cleaned = passes.ring(st).view(np.complex64)[:capture.size].copy()
    cursor.close()
    return cleaned, removed, counters, cleaned_end, cfg
