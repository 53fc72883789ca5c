"""rvx_vrx_batch_noalloc against rvx_vrx_batch, through the C ABI.

The no-allocation entry point runs the same persistent launch with its queue
counter in the library's static scratch. Two calls in a row, the second over
a smaller batch (fewer receivers, shorter spans), must each give the audio
rvx_vrx_batch gives for the same descriptors, bit for bit; a work list the
uint32 queue cannot count is refused before anything launches.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_bit_exact, rng
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_NFM, DEMOD_USB, DEMOD_WFM, GridParams,
                              VrxRequest, build_nco_table, plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 122_880_000
NARROW = GridParams(4096, 16, 2048)
WIDE = GridParams(512, 16, 256)
CAP = 1 << 14
REQUESTS = [(NARROW, VrxRequest(1_000_000, DEMOD_NFM, 12500)),
            (NARROW, VrxRequest(-7_345_678, DEMOD_AM, 10000)),
            (NARROW, VrxRequest(13_000_500, DEMOD_USB, 2700)),
            (WIDE, VrxRequest(20_000_000, DEMOD_WFM, 200_000))]


def _batch(dev, ring_dev, nco_dev, plans, counts_scale):
    from revenant_amd.receivers import Receiver, ReceiverBatch

    rxs = [Receiver(p, (v % 4) * CAP, CAP - 1, ring_dev) for v, p in enumerate(plans)]
    starts, counts = [], []
    for rx in rxs:
        a0, n = rx.audio_range(3 * CAP, 4 * CAP - 1)
        starts.append(a0)
        counts.append(max(1, int(n * counts_scale)))
    audio = [torch.zeros(n * rx.plan.demod.channels, dtype=torch.float32, device="cuda")
             for n, rx in zip(counts, rxs)]
    batch = ReceiverBatch(dev, rxs, nco_dev)
    batch.prepare(starts, counts, audio)
    return batch, audio


def _noalloc(dev, batch):
    return dev.lib.rvx_vrx_batch_noalloc(
        dev.ctx, batch._descs.data_ptr(), batch.n_desc, batch._work.data_ptr(), batch.n_work,
        batch.nco.data_ptr(), batch.nco.numel() // 2, batch.workgroups)


def test_noalloc_batch_matches_the_allocating_one(dev):
    g = rng(0x4E4F414C4C4F4301)
    ring = (g.standard_normal((4, CAP)) + 1j * g.standard_normal((4, CAP))).astype(np.complex64)
    ring_dev = torch.from_numpy(ring.view(np.float32).ravel().copy()).cuda()
    nco_dev = torch.from_numpy(build_nco_table(16).view(np.float32)).cuda()
    plans = [plan_vrx(grid, RATE, req) for grid, req in REQUESTS]
    # two calls in a row, the second over a smaller batch
    for call, (subset, scale) in enumerate(((plans, 1.0), (plans[:2], 0.4))):
        batch, audio = _batch(dev, ring_dev, nco_dev, subset, scale)
        print(f"noalloc call {call}: {len(subset)} receivers, {batch.n_work} items", flush=True)
        batch.launch()  # rvx_vrx_batch, with the caller's counter
        torch.cuda.synchronize()
        want = [a.cpu().numpy().copy() for a in audio]
        for a in audio:
            a.fill_(float("nan"))
        torch.cuda.synchronize()
        assert _noalloc(dev, batch) == 0, _lib.lib().rvx_last_error()
        torch.cuda.synchronize()
        for v, (a, w) in enumerate(zip(audio, want)):
            assert_bit_exact(a.cpu().numpy(), w, what=f"call {call} receiver {v}")


def test_noalloc_batch_refuses_what_the_queue_cannot_count(dev):
    nco_dev = torch.from_numpy(build_nco_table(16).view(np.float32)).cuda()
    dummy = torch.zeros(64, dtype=torch.uint8, device="cuda")
    rc = dev.lib.rvx_vrx_batch_noalloc(dev.ctx, dummy.data_ptr(), 1, dummy.data_ptr(),
                                       0x80000000, nco_dev.data_ptr(), nco_dev.numel() // 2, 0)
    assert rc != 0
    assert b"work items" in _lib.lib().rvx_last_error()
