"""The benchmark's full receiver mix in one persistent launch, every receiver
checked against the oracle bit for bit.

A short span (2^22 samples, 34 ms) of the 122.88 MS/s, 400-carrier cs16
stream: the 4096-channel chain (prescaled cs16 taps) and the 512-channel grid
of the 200 kHz receivers (only their channels stored), then 256 receivers (64
NFM 12.5 kHz, 64 AM 10 kHz, 32 WFM 200 kHz stereo with 75 us de-emphasis, 64
USB 2.7 kHz, 32 CW 500 Hz) in one rvx_vrx_batch launch, replayed from a HIP
graph. Each receiver's audio must equal the oracle's fine stage and
demodulator run on the oracle's own channel rings.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev
from revenant_amd import _lib, dsp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 122_880_000
SAMPLES = 1 << 22
SEED = 0x5743454E45000002
RX_SEED = 0x5652584241544348
NARROW = dsp.GridParams(4096, 16, 2048)
WIDE = dsp.GridParams(512, 16, 256)
FINE_DISPATCH = 8192


def _chain(dev, packed, grid, rows=None):
    proto = dsp.design_prototype(grid, beta=8.0)
    tw = dsp.build_twiddles(grid.channels)
    first = (grid.prototype_length - 1 + grid.decimation - 1) // grid.decimation
    blocks = (SAMPLES - 1) // grid.decimation - first + 1
    ring_blocks = 1 << (blocks - 1).bit_length()
    n_rows = grid.channels if rows is None else len(rows)
    desc = _lib.ChainDesc(dsp.FORMAT_CS16, grid.channels, grid.taps_per_branch, grid.decimation,
                          0, SAMPLES, first, blocks, ring_blocks, 0, np.float32(1 / 16))
    desc.taps_scaled = 1
    keep = None
    if rows is not None:
        table = np.full(grid.channels, 0xFFFFFFFF, dtype=np.uint32)
        for r, k in enumerate(rows):
            table[k] = r
        keep = torch.from_numpy(table.view(np.int32)).cuda()
        desc.ring_rows = len(rows)
        desc.channel_rows = keep.data_ptr()
    ring = torch.zeros(2 * n_rows * ring_blocks, dtype=torch.float32, device="cuda")
    dev.chain(desc, packed, torch.from_numpy(dsp.prescale_taps(proto.taps)).cuda(),
              torch.from_numpy(tw.view(np.float32)).cuda(), ring)
    torch.cuda.synchronize()
    return proto, tw, first, blocks, ring_blocks, ring, keep


def _oracle_audio(plan, chan, chan_base, chan_mask, nco, a0, count):
    from oracle import oracle

    r, hist = plan.demod.decimation, plan.history
    fine_count = (count - 1) * r + hist + 1
    cap = 1 << (fine_count - 1).bit_length()
    f0 = a0 * r - hist
    fine = np.zeros(2 * cap, dtype=np.float32)
    for c in range(0, fine_count, FINE_DISPATCH):
        n = min(FINE_DISPATCH, fine_count - c)
        fp, _ = plan.fine_block(chan_base, chan_mask, cap - 1, f0 + c, n)
        part = oracle.vrx_fine(plan.fine, fp, chan, plan.fine_taps, nco, cap)
        o = (fp.out_offset + np.arange(n)) & (cap - 1)
        fine[2 * o], fine[2 * o + 1] = part[2 * o], part[2 * o + 1]
    dp = plan.demod_block(cap - 1, a0, count)
    return oracle.vrx_demod(plan.demod, dp, fine, plan.weights)


def test_every_receiver_of_the_mix_in_one_launch(dev):
    from oracle import oracle
    from revenant_amd import synth
    from revenant_amd.receivers import Receiver, ReceiverBatch, standard_mix

    print(f"batch256: rendering {SAMPLES} cs16 samples", flush=True)
    scene = synth.wideband_scene(RATE, 400, SEED)
    packed = torch.empty(4 * SAMPLES, dtype=torch.uint8, device="cuda")
    synth.render_device(dev, scene, RATE, 0, SAMPLES, noise_dbfs=-60.0, seed=SEED,
                        fmt=dsp.FORMAT_CS16, out=packed)
    codes = from_dev(packed)

    plans, wide_set = standard_mix(RATE, NARROW, WIDE, RX_SEED)
    assert len(plans) == 256 and len(wide_set) <= 32
    print("batch256: chains", flush=True)
    n_proto, n_tw, n_first, n_blocks, n_rb, n_ring, _ = _chain(dev, packed, NARROW)
    w_proto, w_tw, w_first, w_blocks, w_rb, w_ring, _keep = _chain(dev, packed, WIDE, wide_set)

    # the oracle's rings from the unscaled taps (the prescaled GPU taps must
    # give the same bits)
    od = oracle.OChain(dsp.FORMAT_CS16, 4096, 16, 2048, n_first, n_blocks, n_rb, 0, 0.0, 0.0)
    n_cpu, _ = oracle.chain(od, codes, n_proto.taps, n_tw, threads=8)
    assert_bit_exact(from_dev(n_ring), n_cpu, seed=SEED, what="4096-channel ring")
    od = oracle.OChain(dsp.FORMAT_CS16, 512, 16, 256, w_first, w_blocks, w_rb, 0, 0.0, 0.0)
    w_full, _ = oracle.chain(od, codes, w_proto.taps, w_tw, threads=8)
    w_cpu = w_full.reshape(512, 2 * w_rb)[wide_set].ravel()
    assert_bit_exact(from_dev(w_ring), w_cpu, seed=SEED, what="512-channel selected rows")

    nco_np = dsp.build_nco_table(16)
    nco = torch.from_numpy(nco_np.view(np.float32)).cuda()
    receivers, starts, counts, chans = [], [], [], []
    for name, p in plans:
        if name == "wfm":
            row, rb, first, blocks, ring, cpu = (wide_set.index(p.plan.placement.channel), w_rb,
                                                 w_first, w_blocks, w_ring, w_cpu)
        else:
            row, rb, first, blocks, ring, cpu = (p.plan.placement.channel, n_rb, n_first,
                                                 n_blocks, n_ring, n_cpu)
        rx = Receiver(p, row * rb, rb - 1, ring)
        a0, n = rx.audio_range(first, first + blocks - 1)
        assert n > 0, f"{name} receiver has no audio in the span"
        receivers.append(rx)
        starts.append(a0)
        counts.append(n)
        chans.append(cpu)
    sizes = [n * r.plan.demod.channels for n, r in zip(counts, receivers)]
    audio_buf = torch.zeros(sum(sizes), dtype=torch.float32, device="cuda")
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    audio = [audio_buf[offs[i]:offs[i + 1]] for i in range(len(receivers))]
    batch = ReceiverBatch(dev, receivers, nco)
    batch.prepare(starts, counts, audio)
    print(f"batch256: one launch, {batch.n_desc} segments, {batch.n_work} items", flush=True)
    batch.launch()
    first_run = from_dev(audio_buf).copy()

    # the same launch replayed from a graph (captured on a side stream), the
    # audio poisoned with eager work between replays
    side = torch.cuda.Stream()
    try:
        dev.bind_stream(side)
        with torch.cuda.stream(side):
            dev.graph_begin()
            batch.launch()
            graph = dev.graph_end()
            for rep in range(3):
                audio_buf.fill_(float("nan"))
                side.synchronize()
                graph.launch()
                side.synchronize()
                assert_bit_exact(audio_buf.cpu().numpy(), first_run, what=f"graph replay {rep}")
            graph.close()
    finally:
        dev.bind_stream()

    print("batch256: oracle receivers", flush=True)
    kinds = set()
    for v, ((name, p), rx) in enumerate(zip(plans, receivers)):
        want = _oracle_audio(p, chans[v].view(np.complex64).ravel(), rx.channel_base,
                             rx.channel_mask, nco_np, starts[v], counts[v])
        assert_bit_exact(first_run[offs[v]:offs[v + 1]], want,
                         what=f"receiver {v} ({name}, {p.plan.placement.channel})")
        kinds.add((name, p.demod.channels))
    assert ("wfm", 2) in kinds  # the planner's stereo receivers are in the mix
