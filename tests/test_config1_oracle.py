"""The 2.4 MS/s plain-chain scene through the oracle (CPU): the WFM receiver's
audio against the float64 mix / resample / discriminator, and the tone it
carries.

Pins the composition the GPU test checks bit for bit (channelizer on the
receiver's grid, fine stage, demodulator) against what the mathematics
fixes: a 1 kHz tone at 75 kHz deviation comes out of an FM receiver at the
planner's level, and the float32 path sits within the table quantisation
of the float64 one.
"""
from __future__ import annotations

import numpy as np
import pytest

import config1_scene as c1
from revenant_amd import dsp


@pytest.fixture(scope="module")
def scene():
    from oracle import oracle

    codes = c1.scene_codes()
    grid = c1.receiver_grid()
    proto = dsp.design_prototype(grid, beta=c1.BETA)
    tw = dsp.build_twiddles(grid.channels)
    od, first, blocks, ring_blocks = c1.chain_desc(grid, frame=0)
    ring, _ = oracle.chain(od, codes, proto.taps, tw, threads=4)
    plan = dsp.plan_vrx(grid, c1.RATE, dsp.VrxRequest(c1.WFM_HZ, dsp.DEMOD_WFM, 200_000,
                                                        stereo=False))
    return codes, grid, proto, ring, first, blocks, ring_blocks, plan


def test_receiver_grid_is_what_the_rule_picks(scene):
    _, grid, _, _, _, _, _, plan = scene
    assert (grid.channels, grid.decimation) == (8, 4)
    assert plan.plan.placement.channel == 1 and plan.plan.channel_rate == 600_000


def test_oracle_audio_against_fp64_and_the_tone(scene):
    from oracle import oracle

    codes, grid, proto, ring, first, blocks, ring_blocks, plan = scene
    ch = plan.plan.placement.channel
    f0, count, a0, a_count = c1.receiver_spans(plan, first, blocks)
    a_count = min(a_count, 6000)
    count = min(count, (a0 + a_count) * plan.demod.decimation - f0 + 1)
    nco = dsp.build_nco_table(16)
    fine_cap = 1 << (count + f0).bit_length()
    chan = ring.view(np.complex64).ravel()
    fine = np.zeros(2 * fine_cap, dtype=np.float32)
    for fp in c1.fine_dispatches(plan, ch * ring_blocks, ring_blocks - 1, fine_cap - 1, f0, count):
        part = oracle.vrx_fine(plan.fine, fp, chan, plan.fine_taps, nco, fine_cap)
        o = (fp.out_offset + np.arange(fp.count)) & (fine_cap - 1)
        fine[2 * o], fine[2 * o + 1] = part[2 * o], part[2 * o + 1]
    dp = plan.demod_block(fine_cap - 1, a0, a_count)
    audio = oracle.vrx_demod(plan.demod, dp, fine, plan.weights)

    x = c1.cu8_to_complex(codes)
    y64 = oracle.fp64_channelize(x, proto.taps, grid.channels, grid.decimation, first, blocks)
    ref = c1.fp64_receiver(plan, y64[:, ch], first, f0, count, a0, a_count)
    err = audio.astype(np.float64) - ref
    rms = np.sqrt(np.mean(ref ** 2))
    err_db = 20 * np.log10(np.max(np.abs(err)) / rms)
    print(f"config-1 WFM audio vs float64: max error {err_db:.1f} dB re rms")
    assert err_db < -90.0

    # the 1 kHz tone: 75 kHz deviation reads the planner's full scale
    t = (a0 + np.arange(a_count)) / 48000.0
    settle = 2000
    a = ref[settle:]
    basis = np.exp(-2j * np.pi * 1000.0 * t[settle:])
    tone = 2 * np.abs(np.mean(a * basis))
    resid = a - np.real(2 * np.mean(a * basis) * np.conj(basis))
    assert 20 * np.log10(tone / np.sqrt(np.mean(resid ** 2))) > 20.0
    # full scale is the plan's deviation; the default 75 us de-emphasis
    # takes the analog one-pole's 1/|1 + j 2 pi f tau| at 1 kHz (0.87 dB)
    want_db = -10 * np.log10(1 + (2 * np.pi * 1000.0 * 75e-6) ** 2)
    print(f"config-1 tone {20 * np.log10(tone):.2f} dB, one-pole {want_db:.2f} dB")
    assert abs(20 * np.log10(tone) - want_db) < 0.5
