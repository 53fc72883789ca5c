"""GPU vs oracle, bit for bit: impulse excision's four passes.

After tests/reference/test_impulse_excision.cpp:226-342: three blocks of a
capture with 120 impulses through the oracle to give every side ring a
stream's state, then a fourth block that crosses the ring's end, each pass
compared from identical inputs; then a whole capture through the device and
the oracle, block by block, cleaned ring, removed values and counters
compared.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.excision import (APPLY, FLAG, REFERENCE, SEGMENT, ExcisionCursor,
                                   design_excision, excision_threshold, size_excision_rings)
from tests.excision_stream import OraclePasses, impulsive_capture, run_stream

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 2_400_000


def _pass(cfg, which):
    c = _lib.ExcisionConfig.from_buffer_copy(bytes(cfg))
    c.pass_ = which
    return c


# (source rate, guard override): the narrowband design (hang = lead = 2),
# the wideband one (hang = 12 at 122.88 MS/s) and the widest guards the
# kernels take (hang = lead = 16, the apply pass's LDS window at its largest).
DESIGNS = [(RATE, None), (122_880_000, None), (RATE, 16)]
DESIGN_IDS = ["2.4MS/s", "122.88MS/s", "guard16"]


def _design(rate, guard):
    cfg = design_excision(rate)
    if guard is not None:
        cfg.hang = cfg.lead = guard
    return cfg


@pytest.mark.parametrize("rate,guard", DESIGNS, ids=DESIGN_IDS)
def test_four_passes_match_twins_through_a_wrap(dev, rate, guard):
    wide = rate > RATE
    ring_n, block = (1 << 18, 1 << 15) if wide else (1 << 15, 4096)
    mask = ring_n - 1
    cfg = _design(rate, guard)
    print(f"excision design rate={rate} hang={cfg.hang} lead={cfg.lead} "
          f"seg=2^{cfg.seg_log2} max_width={cfg.max_width}", flush=True)
    rings = size_excision_rings(cfg, block, 1, mask)
    capture = impulsive_capture(ring_n, 0x455843495345, 0.05, 0.0123, 120 * (ring_n >> 15), 6)
    raw = np.ascontiguousarray(capture.view(np.float32).copy())
    power = np.zeros(rings.seg_mask + 1, np.float32)
    level = np.zeros_like(power)
    flags = np.zeros(rings.flag_mask + 1, np.uint32)
    cursor = ExcisionCursor(cfg, mask, rings, excision_threshold(15.0))
    start = ring_n // 2 + 1000
    for at in range(start, start + 3 * block, block):
        warm = cursor.plan(at, block)
        if warm.segment.count:
            oracle.excise_segment(_pass(cfg, SEGMENT), warm.segment, raw, power)
        oracle.excise_reference(_pass(cfg, REFERENCE), warm.reference, power, level)
        oracle.excise_flag(_pass(cfg, FLAG), warm.flag, raw, level, flags)
    plan = cursor.plan(start + 3 * block, block)
    assert plan.segment.count > 0 and plan.apply.count > 0
    assert plan.flag.first + plan.flag.count > ring_n  # crosses the ring's end
    d_raw = to_dev(raw)

    # segment
    want = power.copy()
    oracle.excise_segment(_pass(cfg, SEGMENT), plan.segment, raw, want)
    got = to_dev(power)
    print("excision pass segment", flush=True)
    dev.excise_segment(cfg, plan.segment, d_raw, got)
    assert_bit_exact(from_dev(got), want, what="segment powers")
    power = want
    # reference
    want = level.copy()
    oracle.excise_reference(_pass(cfg, REFERENCE), plan.reference, power, want)
    got = to_dev(level)
    print("excision pass reference", flush=True)
    dev.excise_reference(cfg, plan.reference, to_dev(power), got)
    assert_bit_exact(from_dev(got), want, what="reference levels")
    assert (want > 0).any()
    level = want
    # flag
    want = flags.copy()
    oracle.excise_flag(_pass(cfg, FLAG), plan.flag, raw, level, want)
    got = to_dev(flags.view(np.int32))
    print("excision pass flag", flush=True)
    dev.excise_flag(cfg, plan.flag, d_raw, to_dev(level), got)
    assert np.array_equal(from_dev(got).view(np.uint32), want)
    flags = want
    # apply, over flags carried from earlier blocks; out of place and in place
    w_clean = np.zeros_like(raw)
    w_removed = np.zeros(2 * (rings.removed_mask + 1), np.float32)
    w_counters = np.zeros(4, np.uint32)
    oracle.excise_apply(_pass(cfg, APPLY), plan.apply, raw, flags, w_clean, w_removed, w_counters)
    assert w_counters[1] > 0
    for in_place in (False, True):
        src = to_dev(raw)
        clean = src if in_place else torch.zeros(raw.size, dtype=torch.float32, device="cuda")
        removed = torch.zeros(w_removed.size, dtype=torch.float32, device="cuda")
        counters = torch.zeros(4, dtype=torch.int32, device="cuda")
        print(f"excision pass apply in_place={in_place}", flush=True)
        dev.excise_apply(cfg, plan.apply, src, to_dev(flags.view(np.int32)), clean, removed,
                         counters)
        got_clean = from_dev(clean)
        idx = (plan.apply.first + np.arange(plan.apply.count)) & mask
        sel = np.stack([2 * idx, 2 * idx + 1], 1).ravel()
        assert_bit_exact(got_clean[sel], w_clean[sel], what="cleaned")
        assert_bit_exact(from_dev(removed), w_removed, what="removed")
        assert np.array_equal(from_dev(counters).view(np.uint32), w_counters)
    cursor.close()


class DevicePasses:
    def __init__(self, dev):
        self.dev = dev

    def alloc(self, ring, seg, flags, removed):
        return dict(raw=to_dev(ring.view(np.float32)),
                    power=torch.zeros(seg, dtype=torch.float32, device="cuda"),
                    level=torch.zeros(seg, dtype=torch.float32, device="cuda"),
                    flags=torch.zeros(flags, dtype=torch.int32, device="cuda"),
                    removed=torch.zeros(2 * removed, dtype=torch.float32, device="cuda"),
                    counters=torch.zeros(4, dtype=torch.int32, device="cuda"))

    def run(self, cfg, plan, st):
        d = self.dev
        if plan.segment.count:
            d.excise_segment(cfg, plan.segment, st["raw"], st["power"])
        if plan.reference.count:
            d.excise_reference(cfg, plan.reference, st["power"], st["level"])
        if plan.flag.count:
            d.excise_flag(cfg, plan.flag, st["raw"], st["level"], st["flags"])
        if plan.apply.count:
            st["counters"].zero_()
            d.excise_apply(cfg, plan.apply, st["raw"], st["flags"], st["raw"], st["removed"],
                           st["counters"])
        return from_dev(st["counters"]).view(np.uint32).astype(np.uint64)

    def removed(self, st):
        return from_dev(st["removed"])

    def ring(self, st):
        return from_dev(st["raw"])


@pytest.mark.parametrize("rate,guard", DESIGNS, ids=DESIGN_IDS)
def test_stream_on_device_matches_the_oracle(dev, rate, guard):
    length = 1 << 19
    dirty = impulsive_capture(length, 20261009, 0.1, 0.0371, 40, 4)
    print(f"excision stream on the device rate={rate} guard={guard}", flush=True)
    g_clean, g_removed, g_counters, g_end, _ = run_stream(dirty, rate, 65536, 15.0,
                                                          DevicePasses(dev), guard=guard)
    o_clean, o_removed, o_counters, o_end, _ = run_stream(dirty, rate, 65536, 15.0,
                                                          OraclePasses(), guard=guard)
    assert g_end == o_end
    assert np.array_equal(g_counters, o_counters) and o_counters[1] >= 38
    assert_bit_exact(g_clean, o_clean, what="cleaned stream")
    assert_bit_exact(g_removed[:o_end], o_removed[:o_end], what="removed stream")
