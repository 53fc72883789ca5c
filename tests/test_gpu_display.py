"""GPU: the display tap and the passband pane, after
tests/reference/test_vrx.cpp:1464-1597: the display tap (the fine kernel with
the 256-tap display table) matches its twin bit for bit on every reference
case and on rounding-hostile values, gives the same samples however the
stream is blocked, and the pane (the spectrum kernel over the display ring)
matches its twin and shows a tone outside the receiver's filter at the level
of one inside it, which is what the tap exists for."""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (DEMOD_NFM, VrxRequest, build_nco_table, build_twiddles,
                              plan_vrx, plan_vrx_display)
from tests.display_cases import GRID, RATE, display_cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CHAN_CAP = 1 << 14
CHAN_BASE = CHAN_CAP          # the second of three channels
OUT_CAP = 1 << 12


def _ring(seed, hostile=False):
    g = rng(seed)
    ring = (g.standard_normal((3, CHAN_CAP)) + 1j * g.standard_normal((3, CHAN_CAP)))
    if hostile:
        # magnitudes across many binades and exact halves, where a fused
        # multiply-add or a reordered sum shows
        scale = 2.0 ** g.integers(-20, 20, size=ring.shape)
        ring = np.round(ring * 8) / 8 * scale
    return ring.astype(np.complex64)


def _fine(dev, d, fp, ring, nco):
    out = torch.zeros(2 * OUT_CAP, dtype=torch.float32, device="cuda")
    dev.vrx_fine(d.plan.fine, fp, to_dev(ring.view(np.float32).ravel()),
                 to_dev(d.taps.view(np.float32)), to_dev(nco.view(np.float32)), out)
    return from_dev(out)


@pytest.mark.parametrize("case", display_cases(), ids=lambda c: c[0])
def test_display_tap_matches_twin_on_every_rung(dev, case):
    name, plan, want_r = case
    d = plan_vrx_display(plan)
    assert d.plan.decimation == want_r
    nco = build_nco_table(16)
    ring = _ring(0x5652580000000020)
    fp, _ = d.block(CHAN_BASE, CHAN_CAP - 1, OUT_CAP - 1, 3_000_000_017, 1024)
    cpu = oracle.vrx_fine(d.plan.fine, fp, ring.ravel(), d.taps, nco, OUT_CAP)
    print(f"display tap {name}", flush=True)
    got = _fine(dev, d, fp, ring, nco)
    assert np.count_nonzero(cpu.reshape(-1, 2).any(1)) > 512
    assert_bit_exact(got, cpu, what=f"display {name}")


def test_display_tap_on_rounding_hostile_values(dev):
    plan = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_NFM, 4_000))
    d = plan_vrx_display(plan)
    nco = build_nco_table(16)
    ring = _ring(0x5652580000000021, hostile=True)
    fp, _ = d.block(CHAN_BASE, CHAN_CAP - 1, OUT_CAP - 1, 1, 1000)
    assert fp.in_offset < d.plan.fine.taps
    cpu = oracle.vrx_fine(d.plan.fine, fp, ring.ravel(), d.taps, nco, OUT_CAP)
    print("display tap hostile", flush=True)
    assert_bit_exact(_fine(dev, d, fp, ring, nco), cpu, what="display hostile")


def test_display_tap_is_blocking_invariant(dev):
    _, plan, _ = display_cases()[5]  # cw 500 Hz at 700
    d = plan_vrx_display(plan)
    nco = build_nco_table(16)
    ring = _ring(0x5652580000000022)
    first, total, piece = 7_000_000_003, 900, 300
    fp, _ = d.block(CHAN_BASE, CHAN_CAP - 1, OUT_CAP - 1, first, total)
    whole = _fine(dev, d, fp, ring, nco).reshape(-1, 2)
    for k in range(total // piece):
        fp, _ = d.block(CHAN_BASE, CHAN_CAP - 1, OUT_CAP - 1, first + k * piece, piece)
        part = _fine(dev, d, fp, ring, nco).reshape(-1, 2)
        slots = (fp.out_offset + np.arange(piece)) & (OUT_CAP - 1)
        assert np.array_equal(part[slots].view(np.uint32), whole[slots].view(np.uint32))


def test_the_pane_shows_a_tone_outside_the_filter_at_its_true_level(dev):
    """A 6 kHz NFM receiver (passband +-3 kHz) gets R = 2, a 37.5 kS/s display
    and a +-9.4 kHz pane. Tones of equal amplitude 1 kHz and 6 kHz from the
    receiver's centre read within 1 dB of each other on the pane, and the
    pane is the spectrum twin of the display ring bit for bit."""
    from revenant_amd.delivery import PassbandTap, passband_window

    plan = plan_vrx(GRID, RATE, VrxRequest(196_500, DEMOD_NFM, 6_000))
    tap = PassbandTap(dev, plan, display_ring=1 << 14, transform=1024)
    dp = tap.display.plan
    assert dp.decimation == 2 and dp.display_rate == 37_500
    mix = plan.plan.mix_num / plan.plan.mix_den   # what the fine stage mixes to DC
    fc = plan.plan.channel_rate
    n_in = 1 << 15
    t = np.arange(n_in)
    g = rng(0x50414E45)
    chan = 0.01 * (g.standard_normal(n_in) + 1j * g.standard_normal(n_in))
    for offset in (1_000.0, 6_000.0):
        chan = chan + 0.1 * np.exp(2j * np.pi * (mix + offset) * t / fc)
    ring = chan.astype(np.complex64)
    d_ring = to_dev(ring.view(np.float32))
    count = 8000  # display outputs; their inputs stay inside the ring
    tap.run(d_ring, 0, n_in - 1, 0, count)
    print("passband pane frame", flush=True)
    f = tap.frame(count)
    # the twin of the same frame from the display ring
    disp = from_dev(tap.ring)
    p = _lib.SpectrumParams(2, 1024, 10, 1 << 14, (1 << 14) - 1, (count - 1024) & ((1 << 14) - 1))
    want = oracle.spectrum(p, disp.view(np.complex64), build_twiddles(1024), passband_window(1024))
    assert np.array_equal(f.power_db.view(np.uint32), want[512:].view(np.uint32))
    geo = f.geometry
    levels = []
    for offset in (1_000.0, 6_000.0):
        b = int(round((offset - geo.bin_zero_hz()) / geo.bin_width_hz()))
        levels.append(float(f.power_db[max(0, b - 2):b + 3].max()))
    assert abs(levels[0] - levels[1]) < 1.0, levels
    assert min(levels) > np.median(f.power_db) + 20.0
