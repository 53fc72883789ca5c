"""Host side of the batched receivers: descriptor checks and tile limits.

rvx_vrx_desc_check applies the per-receiver rules of rvx_vrx_fine and
rvx_vrx_demod (core/dsp/vrx_reference.cpp:429-486, vrx_reference.h:555) to a
batch descriptor and refuses tiles that do not fit on chip; no GPU needed.
"""
import ctypes as C

import pytest

from revenant_amd import _lib
from revenant_amd.dsp import DEMOD_AM, DEMOD_NFM, DEMOD_WFM, GridParams, VrxRequest, plan_vrx
from revenant_amd.receivers import MAX_SPAN, max_tile, tile_cost

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


def test_tile_cost_orders_by_work():
    nfm = plan_vrx(GRID, RATE, VrxRequest(1_000_000, DEMOD_NFM, 12500))
    wfm = plan_vrx(WIDE, RATE, VrxRequest(20_000_000, DEMOD_WFM, 200_000))
    assert tile_cost(nfm, 256) < tile_cost(nfm, 512)
    assert tile_cost(wfm, 256) > tile_cost(nfm, 256)
