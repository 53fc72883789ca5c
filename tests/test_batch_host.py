"""Host side of the batched receivers: descriptor checks and tile limits.

rvx_vrx_desc_check applies the per-receiver rules of rvx_vrx_fine and
rvx_vrx_demod (core/dsp/vrx_reference.cpp:429-486, vrx_reference.h:555) to a
batch descriptor and refuses tiles that do not fit on chip; no GPU needed.
"""
import ctypes as C

import pytest

from revenant_amd import _lib
from revenant_amd.dsp import DEMOD_AM, DEMOD_NFM, DEMOD_WFM, GridParams, VrxRequest, plan_vrx
from revenant_amd.receivers import MAX_SPAN, max_tile

RATE = 122_880_000
GRID = GridParams(4096, 16, 2048)
WIDE = GridParams(512, 16, 256)
FAKE = 0x1000  # a non-null stand-in; the check never dereferences pointers


def _desc(plan, first_audio, count, tile):
    r, hist = plan.demod.decimation, plan.history
    d = _lib.VrxDesc()
    d.fine_config = plan.fine
    d.demod_config = plan.demod
    d.fine, _ = plan.fine_block(0, (1 << 16) - 1, 0xFFFFFFFF, first_audio * r - hist,
                                (count - 1) * r + hist + 1)
    d.demod = plan.demod_block(0xFFFFFFFF, first_audio, count)
    d.hist = hist
    d.tile = tile
    d.channel_ring = d.taps = d.weights = d.audio = FAKE
    return d


def _check(d):
    return _lib.lib().rvx_vrx_desc_check(C.byref(d))


@pytest.mark.parametrize("grid,req", [(GRID, VrxRequest(1_000_000, DEMOD_NFM, 12500)),
                                      (GRID, VrxRequest(-7_345_678, DEMOD_AM, 10000)),
                                      (WIDE, VrxRequest(20_000_000, DEMOD_WFM, 200_000))],
                         ids=["nfm", "am", "wfm"])
def test_desc_check_accepts_planned_and_refuses_broken(grid, req):
    plan = plan_vrx(grid, RATE, req)
    t = max_tile(plan)
    assert (t - 1) * plan.demod.decimation + plan.history + 1 <= MAX_SPAN
    assert t * plan.demod.decimation + plan.history + 1 > MAX_SPAN
    good = _desc(plan, 10_000_000, 3000, t)
    assert _check(good) == 0
    too_big = _desc(plan, 10_000_000, 3000, t + 1)
    assert _check(too_big) != 0 and b"fit" in _lib.lib().rvx_last_error()
    wrong_hist = _desc(plan, 10_000_000, 3000, t)
    wrong_hist.hist += 1
    assert _check(wrong_hist) != 0
    short_fine = _desc(plan, 10_000_000, 3000, t)
    short_fine.fine.count -= 1
    assert _check(short_fine) != 0
    no_audio = _desc(plan, 10_000_000, 3000, t)
    no_audio.audio = None
    assert _check(no_audio) != 0
    zero_tile = _desc(plan, 10_000_000, 3000, 0)
    assert _check(zero_tile) != 0


def _span(plan, first_audio, count, tile=0, audio=FAKE):
    return _lib.VrxSpan(C.pointer(plan.plan), 0, (1 << 16) - 1, FAKE, FAKE, FAKE, audio,
                        first_audio, count, tile)


def _plan_batch(spans):
    lib = _lib.lib()
    arr = (_lib.VrxSpan * len(spans))(*spans)
    nd, nw = _lib.u32(), _lib.u32()
    assert lib.rvx_vrx_batch_plan(arr, len(spans), None, 0, C.byref(nd), None, 0,
                                  C.byref(nw)) == 0
    descs, work = (_lib.VrxDesc * nd.value)(), (_lib.VrxWork * nw.value)()
    assert lib.rvx_vrx_batch_plan(arr, len(spans), descs, nd.value, C.byref(nd), work, nw.value,
                                  C.byref(nw)) == 0
    return list(descs), [(w.vrx, w.tile) for w in work]


def test_batch_plan_splits_tiles_and_orders_largest_first():
    """rvx_vrx_batch_plan: every audio output of every span lands in exactly
    one (descriptor, tile) item; a span past the 32-bit resampler recurrence
    is split into descriptors that continue each other; the costliest tiles
    come first."""
    nfm = plan_vrx(GRID, RATE, VrxRequest(1_000_000, DEMOD_NFM, 12500))
    wfm = plan_vrx(WIDE, RATE, VrxRequest(20_000_000, DEMOD_WFM, 200_000))
    spans = [_span(nfm, 5_000_000, 3000, tile=256),
             _span(wfm, 7_000_000, 20000, tile=100, audio=FAKE * 2)]
    descs, work = _plan_batch(spans)
    assert len(descs) > 2  # the WFM span needs several fine blocks
    covered = {}
    for v, d in enumerate(descs):
        assert _lib.lib().rvx_vrx_desc_check(C.byref(d)) == 0
        a0 = (d.fine_config.taps, d.demod.count)
        covered.setdefault(d.fine_config.taps, 0)
        covered[d.fine_config.taps] += d.demod.count
    assert covered == {nfm.fine.taps: 3000, wfm.fine.taps: 20000}
    # WFM descriptors continue each other: audio pointers step by count frames
    # (two floats a frame: the planner's WFM receiver is stereo)
    assert wfm.demod.channels == 2
    w = [d for d in descs if d.fine_config.taps == wfm.fine.taps]
    for a, b in zip(w, w[1:]):
        assert b.audio - a.audio == 4 * a.demod_config.channels * a.demod.count
    # every (descriptor, tile) exactly once
    want = {(v, j) for v, d in enumerate(descs) for j in range(-(-d.demod.count // d.tile))}
    assert sorted(work) == sorted(want)
    # WFM tiles (an 878-tap audio filter and a pilot filter at R = 7) cost more
    # than NFM ones: first
    first_kind = descs[work[0][0]].fine_config.taps
    last_kind = descs[work[-1][0]].fine_config.taps
    assert first_kind == wfm.fine.taps and last_kind == nfm.fine.taps


def test_batch_plan_refuses_reach_below_zero():
    am = plan_vrx(GRID, RATE, VrxRequest(-7_345_678, DEMOD_AM, 10000))
    arr = (_lib.VrxSpan * 1)(_span(am, 3, 100))  # a0 * R < hist
    nd, nw = _lib.u32(), _lib.u32()
    assert _lib.lib().rvx_vrx_batch_plan(arr, 1, None, 0, C.byref(nd), None, 0,
                                         C.byref(nw)) != 0


@pytest.mark.parametrize("mode", [12, 6], ids=["sam", "dsb"])
def test_batch_plan_refuses_carrier_loop_receivers(mode):
    """A SAM or DSB receiver's loop runs over its fine stream in order, so it
    cannot be a batched tile: the planner says so instead of dropping the loop."""
    p = plan_vrx(GRID, RATE, VrxRequest(1_000_000, mode, 6000))
    assert p.plan.carrier.power != 0
    arr = (_lib.VrxSpan * 1)(_span(p, 5_000_000, 1000))
    nd, nw = _lib.u32(), _lib.u32()
    assert _lib.lib().rvx_vrx_batch_plan(arr, 1, None, 0, C.byref(nd), None, 0,
                                         C.byref(nw)) != 0
    assert b"carrier loop" in _lib.lib().rvx_last_error()
