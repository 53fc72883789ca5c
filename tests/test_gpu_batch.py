"""GPU, bit for bit: every receiver of a step in one batched launch.

The batch must equal, to the bit, rvx_vrx_fine followed by rvx_vrx_demod for
each receiver (the per-receiver path the reference records,
core/engine/vrx_stage.cpp:1306-1399), and the oracle's reference_vrx_fine +
reference_vrx_demod composition, whatever the tile split, the number of
persistent workgroups, and across repeated and graph-replayed launches.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_CW, DEMOD_LSB, DEMOD_NFM, DEMOD_RAW, DEMOD_USB,
                              DEMOD_WFM, GridParams, VrxPlan, VrxRequest, build_nco_table,
                              plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 122_880_000
FINE_GRID = GridParams(4096, 16, 2048)
WIDE_GRID = GridParams(512, 16, 256)
REQUESTS = [
    (FINE_GRID, VrxRequest(1_000_000, DEMOD_NFM, 12500)),
    (FINE_GRID, VrxRequest(-7_345_678, DEMOD_AM, 10000)),
    (FINE_GRID, VrxRequest(13_000_500, DEMOD_USB, 2700)),
    (FINE_GRID, VrxRequest(-20_000_000, DEMOD_LSB, 2700)),
    (FINE_GRID, VrxRequest(33_333_333, DEMOD_CW, 500, cw_pitch=700)),
    (FINE_GRID, VrxRequest(5_000_000, DEMOD_RAW, 12000)),
    (WIDE_GRID, VrxRequest(20_000_000, DEMOD_WFM, 200_000, stereo=False)),
]
CAP = 1 << 14          # channel length in the test ring
CHANNELS = 8


def stereo_plan(base: VrxPlan, seed: int) -> VrxPlan:
    """The mono WFM plan turned into a two-channel receiver with a short pilot
    filter and seeded weights (every weight of the stereo branch exercised)."""
    plan = type(base.plan).from_buffer_copy(base.plan)
    plan.demod.channels = 2
    plan.demod.pilot_taps = 33
    d = plan.demod
    n = d.audio_taps + d.dc_taps + 2 * d.pilot_taps + 2 * d.audio_taps
    w = rng(seed).uniform(-0.05, 0.05, n).astype(np.float32)
    w[:d.audio_taps] = base.weights[:d.audio_taps]
    return VrxPlan(plan, base.fine_taps, w)


@pytest.fixture(scope="module")
def scene():
    nco = build_nco_table(16)
    g = rng(0x5652580000000C00)
    ring = (g.standard_normal((CHANNELS, CAP)) + 1j * g.standard_normal((CHANNELS, CAP))).astype(
        np.complex64)
    plans = [plan_vrx(grid, RATE, req) for grid, req in REQUESTS]
    plans.append(stereo_plan(plans[-1], 0x5652580000000C01))
    # the planner's own stereo receiver: pilot bandpass, rotation and 75 us curve
    plans.append(plan_vrx(WIDE_GRID, RATE, VrxRequest(-20_000_000, DEMOD_WFM, 200_000)))
    return nco, ring, plans


def _receivers(dev, scene, ring_dev):
    from revenant_amd.receivers import Receiver

    nco, ring, plans = scene
    return [Receiver(p, (v % CHANNELS) * CAP, CAP - 1, ring_dev) for v, p in enumerate(plans)]


def _oracle_audio(rx, ring, nco, first_audio, count):
    p = rx.plan
    r, hist = p.demod.decimation, p.history
    fine_count = (count - 1) * r + hist + 1
    out_cap = 1 << int(np.ceil(np.log2(fine_count)))
    f0 = first_audio * r - hist
    fp, _ = p.fine_block(rx.channel_base, rx.channel_mask, out_cap - 1, f0, fine_count)
    fine = oracle.vrx_fine(p.fine, fp, ring.ravel(), p.fine_taps, nco, out_cap)
    dp = p.demod_block(out_cap - 1, first_audio, count)
    return oracle.vrx_demod(p.demod, dp, fine, p.weights)


def _per_receiver_audio(dev, rx, nco_dev, first_audio, count):
    p = rx.plan
    r, hist = p.demod.decimation, p.history
    fine_count = (count - 1) * r + hist + 1
    out_cap = 1 << int(np.ceil(np.log2(fine_count)))
    f0 = first_audio * r - hist
    fp, _ = p.fine_block(rx.channel_base, rx.channel_mask, out_cap - 1, f0, fine_count)
    fine = torch.zeros(2 * out_cap, dtype=torch.float32, device="cuda")
    dev.vrx_fine(p.fine, fp, rx.ring, rx.taps, nco_dev, fine)
    dp = p.demod_block(out_cap - 1, first_audio, count)
    audio = torch.zeros(count * p.demod.channels, dtype=torch.float32, device="cuda")
    dev.vrx_demod(p.demod, dp, fine, rx.weights, audio)
    return from_dev(audio)


def _starts(plans, hours=3):
    # far into the stream, so 64-bit indices and u32 wraps are exercised
    return [hours * 3600 * int(p.plan.audio_rate) + 17 * v for v, p in enumerate(plans)]


@pytest.mark.parametrize("workgroups", [0, 3])
@pytest.mark.parametrize("tile", [None, 1, 100, 512])
def test_batch_matches_per_receiver_and_oracle(dev, scene, tile, workgroups):
    from revenant_amd.receivers import ReceiverBatch, max_tile

    nco, ring, plans = scene
    ring_dev = to_dev(ring.view(np.float32).ravel())
    nco_dev = to_dev(nco.view(np.float32))
    rxs = _receivers(dev, scene, ring_dev)
    batch = ReceiverBatch(dev, rxs, nco_dev, workgroups=workgroups)
    starts = _starts(plans)
    counts = [1500 + 37 * v for v in range(len(rxs))]
    audio = [torch.full((n * r.plan.demod.channels,), float("nan"), device="cuda")
             for n, r in zip(counts, rxs)]
    tiles = None if tile is None else [min(tile, max_tile(r.plan)) for r in rxs]
    batch.prepare(starts, counts, audio, tiles)
    batch.launch()
    for v, rx in enumerate(rxs):
        got = from_dev(audio[v])
        ref = _per_receiver_audio(dev, rx, nco_dev, starts[v], counts[v])
        assert_bit_exact(got, ref, what=f"batch vs per-receiver, receiver {v} tile {tile}")
        if tile is None and workgroups == 0:
            cpu = _oracle_audio(rx, ring, nco, starts[v], counts[v])
            assert_bit_exact(got, cpu, what=f"batch vs oracle, receiver {v}")


SMALL = 1 << 10  # a channel ring the fine windows cross every pass


@pytest.mark.parametrize("tile", [None, 64])
def test_batch_windows_across_the_channel_ring_wrap(dev, scene, tile):
    """On a 1024-sample channel ring the fine windows (up to 256 taps) cross
    the ring's wrap many times per receiver, so the fine loop's wrapping form
    runs beside its straight one (which needs a window clear of the wrap);
    both against the oracle, bit for bit."""
    from revenant_amd.receivers import Receiver, ReceiverBatch, max_tile

    nco, ring, plans = scene
    small = np.ascontiguousarray(ring[:, :SMALL])
    ring_dev = to_dev(small.view(np.float32).ravel())
    nco_dev = to_dev(nco.view(np.float32))
    rxs = [Receiver(p, (v % CHANNELS) * SMALL, SMALL - 1, ring_dev) for v, p in enumerate(plans)]
    batch = ReceiverBatch(dev, rxs, nco_dev)
    starts = _starts(plans)
    counts = [900 + 29 * v for v in range(len(rxs))]
    audio = [torch.full((n * r.plan.demod.channels,), float("nan"), device="cuda")
             for n, r in zip(counts, rxs)]
    tiles = None if tile is None else [min(tile, max_tile(r.plan)) for r in rxs]
    batch.prepare(starts, counts, audio, tiles)
    batch.launch()
    for v, rx in enumerate(rxs):
        cpu = _oracle_audio(rx, small, nco, starts[v], counts[v])
        assert_bit_exact(from_dev(audio[v]), cpu, what=f"batch vs oracle across the wrap, receiver {v}")


def test_batch_relaunch_and_graph_replay(dev, scene):
    from revenant_amd.receivers import ReceiverBatch

    nco, ring, plans = scene
    side = torch.cuda.Stream()
    try:
        dev.bind_stream(side)
        with torch.cuda.stream(side):
            ring_dev = to_dev(ring.view(np.float32).ravel())
            nco_dev = to_dev(nco.view(np.float32))
            rxs = _receivers(dev, scene, ring_dev)
            batch = ReceiverBatch(dev, rxs, nco_dev)
            starts = _starts(plans, hours=1)
            counts = [2048] * len(rxs)
            audio = [torch.zeros(n * r.plan.demod.channels, device="cuda")
                     for n, r in zip(counts, rxs)]
            batch.prepare(starts, counts, audio)
            batch.launch()
            side.synchronize()
            eager = [a.cpu().numpy().copy() for a in audio]
            # a second launch re-zeroes its own queue counter
            for a in audio:
                a.fill_(float("nan"))
            batch.launch()
            side.synchronize()
            for v, a in enumerate(audio):
                assert_bit_exact(a.cpu().numpy(), eager[v], what=f"relaunch receiver {v}")
            dev.graph_begin()
            batch.launch()
            graph = dev.graph_end()
            for replay in range(3):
                for a in audio:
                    a.fill_(float("nan"))
                side.synchronize()
                graph.launch()
                side.synchronize()
                for v, a in enumerate(audio):
                    assert_bit_exact(a.cpu().numpy(), eager[v],
                                     what=f"graph replay {replay} receiver {v}")
            graph.close()
    finally:
        torch.cuda.synchronize()
        dev.bind_stream()


def test_batch_tuner_keeps_bits(dev, scene):
    from revenant_amd.receivers import ReceiverBatch

    nco, ring, plans = scene
    ring_dev = to_dev(ring.view(np.float32).ravel())
    nco_dev = to_dev(nco.view(np.float32))
    rxs = _receivers(dev, scene, ring_dev)
    batch = ReceiverBatch(dev, rxs, nco_dev)
    starts = _starts(plans)
    counts = [4000] * len(rxs)
    audio = [torch.zeros(n * r.plan.demod.channels, device="cuda") for n, r in zip(counts, rxs)]
    batch.prepare(starts, counts, audio)
    batch.launch()
    ref = [from_dev(a).copy() for a in audio]
    tiles = batch.tune(starts, counts, audio, candidates=(64, 256, 1024), reps=1)
    assert len(tiles) == len(rxs)
    for a in audio:
        a.zero_()
    batch.launch()
    for v, a in enumerate(audio):
        assert_bit_exact(from_dev(a), ref[v], what=f"tuned tiles {tiles} receiver {v}")


def test_batch_splits_long_blocks(dev, scene):
    """A block longer than one fine block's 32-bit recurrence allows is split
    into several descriptors; outputs are pure in the absolute index, so the
    bits equal the per-receiver path run in chunks of another size."""
    from revenant_amd.receivers import ReceiverBatch

    nco, ring, plans = scene
    ring_dev = to_dev(ring.view(np.float32).ravel())
    nco_dev = to_dev(nco.view(np.float32))
    rxs = _receivers(dev, scene, ring_dev)[-2:]  # WFM mono and stereo, R = 7
    batch = ReceiverBatch(dev, rxs, nco_dev)
    count = 20000
    starts = [12345678, 23456789]
    audio = [torch.zeros(count * r.plan.demod.channels, device="cuda") for r in rxs]
    batch.prepare(starts, [count] * 2, audio)
    assert batch.n_desc > 2
    batch.launch()
    for v, rx in enumerate(rxs):
        ch = rx.plan.demod.channels
        ref = np.concatenate([
            _per_receiver_audio(dev, rx, nco_dev, starts[v] + lo, min(2500, count - lo))
            for lo in range(0, count, 2500)])
        assert_bit_exact(from_dev(audio[v]), ref[:count * ch], what=f"split receiver {v}")


def test_unbatched_path_gives_the_same_bits(dev, scene):
    """The per-segment fine + demod launches (the reference's recording) write
    exactly the batch's audio."""
    from revenant_amd.receivers import ReceiverBatch

    nco, ring, plans = scene
    ring_dev = to_dev(ring.view(np.float32).ravel())
    nco_dev = to_dev(nco.view(np.float32))
    rxs = _receivers(dev, scene, ring_dev)
    batch = ReceiverBatch(dev, rxs, nco_dev)
    starts = _starts(plans)
    counts = [3000] * len(rxs)
    audio = [torch.zeros(n * r.plan.demod.channels, device="cuda") for n, r in zip(counts, rxs)]
    batch.prepare(starts, counts, audio)
    batch.launch()
    ref = [from_dev(a).copy() for a in audio]
    for a in audio:
        a.fill_(float("nan"))
    batch.unbatched(torch.zeros(2 << 15, dtype=torch.float32, device="cuda"))
    for v, a in enumerate(audio):
        assert_bit_exact(from_dev(a), ref[v], what=f"unbatched receiver {v}")
