"""Pins the receiver oracle (fine stage + demodulators) and the planner.

After tests/reference/test_vrx.cpp: blocking invariance (:1241), three hours
into a stream (:1291), a unit tone at the centre (:1951) and demodulators
recovering their modulation (:2156).
"""
import numpy as np
import pytest

from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_AM, DEMOD_CW, DEMOD_LSB, DEMOD_NFM, DEMOD_USB, DEMOD_WFM,
                              GridParams, VrxRequest, build_nco_table, nco_delta, plan_vrx)

GRID = GridParams(64, 16, 32)        # 2.4 MS/s -> 75 kS/s channels
RATE = 2_400_000
NCO = build_nco_table(16)


def _fine(plan, ring, first_output, count, chan_mask, out_cap=1 << 16):
    fp, first_input = plan.fine_block(0, chan_mask, out_cap - 1, first_output, count)
    out = oracle.vrx_fine(plan.fine, fp, ring, plan.fine_taps, NCO, out_cap)
    return out.view(np.complex64), fp, first_input


def test_planner_rates_and_placement():
    p = plan_vrx(GridParams(4096, 16, 2048), 122_880_000, VrxRequest(1_000_000, DEMOD_NFM, 12500))
    pl = p.plan
    assert pl.placement.channel == round(1_000_000 * 4096 / 122_880_000)
    # residual is exact: centre minus channel centre, in (-rate/2, rate/2]
    centre = pl.placement.centre_num / pl.placement.centre_den
    assert pl.placement.residual_num / pl.placement.residual_den == 1_000_000 - centre
    assert pl.demod_rate % 48000 == 0 and pl.demod_rate >= 18750
    assert pl.demod.decimation == pl.demod_rate // 48000
    # P*T taps plus the one zero entry the lerp of the last phase reads
    assert p.fine_taps.size == pl.fine.phases * pl.fine.taps + 1 and p.fine_taps[-1] == 0
    assert pl.fine_nco_delta == nco_delta(pl.mix_num, pl.mix_den, pl.demod_rate)


def test_planner_refuses_an_unplaceable_band():
    with pytest.raises(_lib.RvxError):
        plan_vrx(GridParams(4096, 16, 2048), 122_880_000, VrxRequest(0, DEMOD_WFM, 200_000))


@pytest.mark.parametrize("mode,bw", [(DEMOD_NFM, 12500), (DEMOD_AM, 10000), (DEMOD_USB, 2700),
                                     (DEMOD_CW, 500)])
def test_fine_blocking_invariance(mode, bw):
    """Two dispatches of n outputs equal one of 2n, bit for bit."""
    p = plan_vrx(GRID, RATE, VrxRequest(310_000, mode, bw, cw_pitch=700))
    cap = 1 << 15
    g = np.random.default_rng(mode)
    ring = (g.standard_normal(cap) + 1j * g.standard_normal(cap)).astype(np.complex64)
    j0, n = 12345, 600
    whole, _, _ = _fine(p, ring, j0, 2 * n, cap - 1)
    a, _, _ = _fine(p, ring, j0, n, cap - 1)
    b, _, _ = _fine(p, ring, j0 + n, n, cap - 1)
    idx = (np.arange(2 * n) + j0) & 0xFFFF
    got = np.concatenate([a[idx[:n]], b[idx[n:]]])
    assert np.array_equal(got.view(np.uint32), whole[idx].view(np.uint32))


def test_fine_three_hours_in():
    """An absolute index past 2^32 fine-rate phase turns: same bits whichever
    block boundary a dispatch starts on."""
    p = plan_vrx(GRID, RATE, VrxRequest(-510_000, DEMOD_NFM, 12500))
    cap = 1 << 15
    g = np.random.default_rng(3)
    ring = (g.standard_normal(cap) + 1j * g.standard_normal(cap)).astype(np.complex64)
    j0 = 3 * 3600 * p.plan.demod_rate + 17
    whole, fp, _ = _fine(p, ring, j0, 512, cap - 1)
    a, _, _ = _fine(p, ring, j0 + 256, 256, cap - 1)
    idx = (np.arange(512) + j0) & 0xFFFF
    assert np.array_equal(whole[idx[256:]].view(np.uint32), a[idx[256:]].view(np.uint32))
    # the 0.64 phase of the first output is delta * j0 mod 2^64
    phase = (p.plan.fine_nco_delta * j0) % (1 << 64)
    assert (fp.nco_phase_high << 32 | fp.nco_phase_low) == phase


def test_unit_tone_at_the_centre():
    """A tone at the receiver's centre comes out of the fine stage at DC,
    unit level (vrx filter unit gain, exact NCO)."""
    p = plan_vrx(GRID, RATE, VrxRequest(150_000, DEMOD_NFM, 12500))
    pl = p.plan
    fc = pl.channel_rate
    resid = pl.placement.residual_num / pl.placement.residual_den
    cap = 1 << 15
    n = np.arange(cap)
    ring = np.exp(2j * np.pi * resid * n / fc).astype(np.complex64)
    out, fp, first_input = _fine(p, ring, 2000, 1000, cap - 1)
    y = out[(np.arange(1000) + 2000) & 0xFFFF]
    assert np.max(np.abs(np.abs(y) - 1.0)) < 2e-3
    # at DC: the phase does not advance
    dphi = np.angle(y[1:] * np.conj(y[:-1]))
    assert np.max(np.abs(dphi)) < 1e-3


def _demod(p, fine, count, first_audio=0):
    cap = fine.size
    dp = p.demod_block(cap - 1, first_audio, count)
    return oracle.vrx_demod(p.demod, dp, fine, p.weights)


def test_fm_discriminator_reads_deviation():
    """A carrier offset by f from the centre reads f / deviation at the output."""
    p = plan_vrx(GRID, RATE, VrxRequest(0, DEMOD_NFM, 12500))
    fd, dev = p.plan.demod_rate, p.plan.deviation
    cap = 1 << 14
    f = dev / 2
    fine = np.exp(2j * np.pi * f * np.arange(cap) / fd).astype(np.complex64)
    audio = _demod(p, fine, 500, first_audio=200)
    assert np.allclose(audio, 0.5, atol=2e-4)


def test_am_recovers_modulation():
    p = plan_vrx(GRID, RATE, VrxRequest(0, DEMOD_AM, 10000))
    fd = p.plan.demod_rate
    cap = 1 << 15
    n = np.arange(cap)
    m, fm = 0.5, 1000.0
    fine = ((1 + m * np.cos(2 * np.pi * fm * n / fd)) * np.exp(1j * 0.3)).astype(np.complex64)
    R = p.plan.demod.decimation
    audio = _demod(p, fine, 4000, first_audio=2000)
    assert abs(audio.max() - m) < 0.02 and abs(audio.min() + m) < 0.02


@pytest.mark.parametrize("mode,sign", [(DEMOD_USB, 1), (DEMOD_LSB, -1)])
def test_ssb_detects_real_part(mode, sign):
    p = plan_vrx(GRID, RATE, VrxRequest(0, mode, 2700))
    cap = 1 << 13
    fine = (np.arange(cap) % 7 + 1j * 5).astype(np.complex64)
    dp = p.demod_block(cap - 1, 300, 100)
    audio = oracle.vrx_demod(p.demod, dp, fine, p.weights)
    # audio = gain * sum_t re(x[idx - t]) w[t]
    idx = 300 * p.demod.decimation + np.arange(100) * p.demod.decimation
    w = p.weights[:p.demod.audio_taps].astype(np.float64)
    want = np.array([sum(float(fine[(i - t) % cap].real) * w[t] for t in range(w.size))
                     for i in idx]) * p.plan.demod_gain
    assert np.allclose(audio, want, atol=1e-5)
