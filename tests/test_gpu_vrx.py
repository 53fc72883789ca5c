"""GPU vs oracle, bit for bit: receiver fine stage and demodulators.

After tests/reference/test_vrx.cpp:1206-1690 (conformance of vrx_fine and
vrx_demod against their twins), with the receivers of the batching workload:
NFM, AM, WFM (75 us de-emphasis folded in, 50 us, and none), USB, LSB, CW
and a raw tap.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEEMPH_EU50, DEEMPH_NONE, DEMOD_AM, DEMOD_CW, DEMOD_LSB,
                              DEMOD_NFM, DEMOD_RAW, DEMOD_USB, DEMOD_WFM, GridParams, VrxRequest,
                              build_nco_table, plan_vrx)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 122_880_000
FINE_GRID = GridParams(4096, 16, 2048)
WIDE_GRID = GridParams(512, 16, 256)
CASES = [
    (FINE_GRID, VrxRequest(1_000_000, DEMOD_NFM, 12500)),
    (FINE_GRID, VrxRequest(-7_345_678, DEMOD_AM, 10000)),
    (FINE_GRID, VrxRequest(13_000_500, DEMOD_USB, 2700)),
    (FINE_GRID, VrxRequest(-20_000_000, DEMOD_LSB, 2700)),
    (FINE_GRID, VrxRequest(33_333_333, DEMOD_CW, 500, cw_pitch=700)),
    (FINE_GRID, VrxRequest(5_000_000, DEMOD_RAW, 12000)),
    (WIDE_GRID, VrxRequest(20_000_000, DEMOD_WFM, 200_000)),
    (WIDE_GRID, VrxRequest(20_000_000, DEMOD_WFM, 200_000, deemphasis=DEEMPH_EU50)),
    (WIDE_GRID, VrxRequest(20_000_000, DEMOD_WFM, 200_000, deemphasis=DEEMPH_NONE)),
]
IDS = ["nfm", "am", "usb", "lsb", "cw", "raw", "wfm", "wfm-50us", "wfm-flat"]


@pytest.fixture(scope="module")
def nco():
    return build_nco_table(16)


@pytest.mark.parametrize("grid,req", CASES, ids=IDS)
def test_fine_matches_twin(dev, nco, grid, req):
    p = plan_vrx(grid, RATE, req)
    cap = 1 << 14
    g = rng(0x56525800000000A0 + req.mode)
    chans = 3
    ring = (g.standard_normal((chans, cap)) + 1j * g.standard_normal((chans, cap))).astype(
        np.complex64)
    chan_base = cap  # second channel of the ring
    out_cap = 1 << 13
    first_output = 3 * 3600 * p.plan.demod_rate + 11  # three hours in
    count = 3000
    fp, _ = p.fine_block(chan_base, cap - 1, out_cap - 1, first_output, count)
    cpu = oracle.vrx_fine(p.fine, fp, ring.ravel(), p.fine_taps, nco, out_cap)
    out = torch.zeros(2 * out_cap, dtype=torch.float32, device="cuda")
    dev.vrx_fine(p.fine, fp, to_dev(ring.view(np.float32).ravel()),
                 to_dev(p.fine_taps.view(np.float32)), to_dev(nco.view(np.float32)), out)
    assert_bit_exact(from_dev(out), cpu, what=f"vrx_fine {IDS[CASES.index((grid, req))]}")


@pytest.mark.parametrize("grid,req", CASES, ids=IDS)
@pytest.mark.parametrize("count", [1, 777, 5000])
def test_demod_matches_twin(dev, grid, req, count):
    p = plan_vrx(grid, RATE, req)
    g = rng(0x56525800000000B0 + req.mode)
    cap = 1 << 16
    t = np.arange(cap)
    # a plausible baseband: a wandering carrier plus noise
    fine = (np.exp(1j * (0.01 * t + 2.0 * np.sin(t / 300.0))) * (1 + 0.5 * np.cos(t / 50.0)) +
            0.05 * (g.standard_normal(cap) + 1j * g.standard_normal(cap))).astype(np.complex64)
    first_audio = (1 << 31) // max(1, p.demod.decimation) + 5  # offset past the 32-bit seam
    dp = p.demod_block(cap - 1, first_audio, count)
    cpu = oracle.vrx_demod(p.demod, dp, fine, p.weights)
    audio = torch.zeros(count * p.demod.channels, dtype=torch.float32, device="cuda")
    dev.vrx_demod(p.demod, dp, to_dev(fine.view(np.float32)), to_dev(p.weights), audio)
    assert_bit_exact(from_dev(audio), cpu, what=f"vrx_demod {req.mode} count {count}")


def test_demod_refuses_short_ring(dev):
    p = plan_vrx(FINE_GRID, RATE, VrxRequest(-7_345_678, DEMOD_AM, 10000))
    fine = torch.zeros(2 * 256, dtype=torch.float32, device="cuda")
    audio = torch.zeros(4096, dtype=torch.float32, device="cuda")
    dp = _lib.VrxDemodParams(255, 0, 4096, 1.0)
    with pytest.raises(_lib.RvxError):
        dev.vrx_demod(p.demod, dp, fine, to_dev(p.weights), audio)


def test_stereo_station_matches_twin_and_separates(dev):
    """A pilot-tone stereo multiplex through the planner's stereo receiver:
    bit-exact with the oracle, and the two tones land in their own channels."""
    import test_stereo as st

    p = st._wfm(deemphasis=DEEMPH_NONE)
    ring = st._station(p, 1000.0, 3000.0)
    cap, R = ring.size, p.demod.decimation
    first = -(-p.history // R) + 8
    count = (cap - 1) // R - first
    dp = p.demod_block(cap - 1, first, count)
    cpu = oracle.vrx_demod(p.demod, dp, ring, p.weights)
    audio = torch.zeros(2 * count, dtype=torch.float32, device="cuda")
    dev.vrx_demod(p.demod, dp, to_dev(ring.view(np.float32)), to_dev(p.weights), audio)
    got = from_dev(audio)
    assert_bit_exact(got, cpu, what="vrx_demod stereo station")
    t = (first + np.arange(count)) / float(p.plan.output_rate)
    l_want, l_leak = st._tones(got[0::2], t, (1000.0, 3000.0))
    r_leak, r_want = st._tones(got[1::2], t, (1000.0, 3000.0))
    assert 20 * np.log10(l_want / l_leak) > 30.0 and 20 * np.log10(r_want / r_leak) > 30.0
