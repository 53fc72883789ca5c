"""The 2.4 MS/s plain-chain scene end to end on the GPU, bit for bit.

2^20 cu8 samples (WFM at +200 kHz, 75 kHz deviation, 1 kHz tone, AWGN 20 dB
below): the 64-channel critically sampled chain (16 taps per branch, Kaiser
beta 8) over every block with full history, and the 200 kHz WFM receiver on
the 8-channel grid channel_count_for picks for it (core/engine/vrx_place.cpp:187)
from the same codes: chain, fine stage (dispatches of 8192 outputs) and
demodulator (the planner's default stereo receiver, and mono), each compared
with the oracle bit for bit. The mono audio's error against the float64
receiver is reported (the oracle's own is pinned in test_config1_oracle.py).
"""
from __future__ import annotations

import numpy as np
import pytest

import config1_scene as c1
from helpers import assert_bit_exact, from_dev, to_dev
from revenant_amd import _lib, dsp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def codes():
    return c1.scene_codes()


def _chain_both(dev, codes, grid, frame):
    from oracle import oracle

    proto = dsp.design_prototype(grid, beta=c1.BETA)
    tw = dsp.build_twiddles(grid.channels)
    od, first, blocks, ring_blocks = c1.chain_desc(grid, frame=frame)
    alpha = np.float32(1 / 16)
    desc = _lib.ChainDesc(dsp.FORMAT_CU8, grid.channels, grid.taps_per_branch, grid.decimation,
                          0, c1.SAMPLES, first, blocks, ring_blocks, frame, alpha)
    ring = torch.zeros(2 * grid.channels * ring_blocks, dtype=torch.float32, device="cuda")
    partial = torch.zeros(max(1, blocks // max(frame, 1)) * grid.channels, dtype=torch.float32,
                          device="cuda")
    print(f"config-1 chain M={grid.channels} D={grid.decimation} blocks={blocks}", flush=True)
    dev.chain(desc, to_dev(codes), to_dev(proto.taps), to_dev(tw.view(np.float32)), ring,
              partial if frame else None)
    ring_cpu, part_cpu = oracle.chain(od, codes, proto.taps, tw, threads=8)
    return proto, ring, from_dev(ring), ring_cpu, from_dev(partial), part_cpu, first, blocks, \
        ring_blocks


def test_plain_chain_every_block_bit_exact(dev, codes):
    _, _, ring, ring_cpu, part, part_cpu, _, blocks, _ = _chain_both(dev, codes, c1.PLAIN, 16)
    assert blocks == (1 << 20) // 64 - 16
    assert_bit_exact(ring, ring_cpu, seed=c1.SEED, what="config-1 channel ring")
    assert_bit_exact(part, part_cpu, seed=c1.SEED, what="config-1 averaged spectrum")


def test_wfm_receiver_on_its_grid_bit_exact(dev, codes):
    from oracle import oracle

    grid = c1.receiver_grid()
    assert (grid.channels, grid.decimation) == (8, 4)
    proto, ring_dev, ring, ring_cpu, part, part_cpu, first, blocks, ring_blocks = _chain_both(
        dev, codes, grid, 16)
    assert_bit_exact(ring, ring_cpu, seed=c1.SEED, what="receiver grid channel ring")
    assert_bit_exact(part, part_cpu, seed=c1.SEED, what="receiver grid averaged spectrum")

    nco = dsp.build_nco_table(16)
    nco_dev = to_dev(nco.view(np.float32))
    chan = ring_cpu.view(np.complex64).ravel()
    for stereo in (True, False):
        plan = dsp.plan_vrx(grid, c1.RATE,
                            dsp.VrxRequest(c1.WFM_HZ, dsp.DEMOD_WFM, 200_000, stereo=stereo))
        ch = plan.plan.placement.channel
        f0, count, a0, a_count = c1.receiver_spans(plan, first, blocks)
        fine_cap = 1 << (f0 + count).bit_length()
        fine_dev = torch.zeros(2 * fine_cap, dtype=torch.float32, device="cuda")
        fine_cpu = np.zeros(2 * fine_cap, dtype=np.float32)
        taps_dev = to_dev(plan.fine_taps.view(np.float32))
        fps = c1.fine_dispatches(plan, ch * ring_blocks, ring_blocks - 1, fine_cap - 1, f0, count)
        print(f"config-1 receiver stereo={stereo} fine={count} in {len(fps)} dispatches, "
              f"audio={a_count}", flush=True)
        for fp in fps:
            dev.vrx_fine(plan.fine, fp, ring_dev, taps_dev, nco_dev, fine_dev)
            part_fine = oracle.vrx_fine(plan.fine, fp, chan, plan.fine_taps, nco, fine_cap)
            o = (fp.out_offset + np.arange(fp.count)) & (fine_cap - 1)
            fine_cpu[2 * o], fine_cpu[2 * o + 1] = part_fine[2 * o], part_fine[2 * o + 1]
        assert_bit_exact(from_dev(fine_dev), fine_cpu, seed=c1.SEED, what="config-1 fine ring")

        dp = plan.demod_block(fine_cap - 1, a0, a_count)
        audio = torch.zeros(a_count * plan.demod.channels, dtype=torch.float32, device="cuda")
        dev.vrx_demod(plan.demod, dp, fine_dev, to_dev(plan.weights), audio)
        got = from_dev(audio)
        assert_bit_exact(got, oracle.vrx_demod(plan.demod, dp, fine_cpu, plan.weights),
                         seed=c1.SEED, what=f"config-1 audio stereo={stereo}")
        if not stereo:
            x = c1.cu8_to_complex(codes)
            y64 = oracle.fp64_channelize(x, proto.taps, grid.channels, grid.decimation, first,
                                         blocks)
            n = min(a_count, 6000)
            ref = c1.fp64_receiver(plan, y64[:, ch], first, f0, count, a0, n)
            err = got[:n].astype(np.float64) - ref
            err_db = 20 * np.log10(np.max(np.abs(err)) / np.sqrt(np.mean(ref ** 2)))
            print(f"config-1 WFM audio (GPU) vs float64: max error {err_db:.1f} dB re rms",
                  flush=True)
            assert err_db < -90.0
