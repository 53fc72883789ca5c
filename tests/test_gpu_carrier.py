"""GPU vs oracle, bit for bit: the carrier loop.

After tests/reference/test_vrx_carrier.cpp:420-526: a stream that makes the
loop do everything (noise, an offset carrier, a fade, a frequency step) in
dispatches of awkward lengths starting part way round the ring so it wraps,
a restart at the fifth and a retune shift at the seventh; ring and state
compared after every dispatch, and the loop locked by the end so the diff
compared a trajectory. Then the batched launch: many receivers of both loops,
ring sizes and dispatch lengths, one lane each, against the oracle per
receiver over several launches.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import DEMOD_DSB, DEMOD_SAM
from tests.carrier_signals import TWO_PI, busy_stream, dsb_plan, nco_for, sam_plan

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LENGTHS = [1, 7, 64, 333, 1000, 2048, 5, 900, 1642]


@pytest.mark.parametrize("mode,seed", [(DEMOD_SAM, 0x53414D0000000001),
                                       (DEMOD_DSB, 0x4453420000000001)], ids=["sam", "dsb"])
def test_carrier_matches_twin_across_dispatch_boundaries(dev, mode, seed):
    plan = sam_plan() if mode == DEMOD_SAM else dsb_plan()
    rate = float(plan.plan.demod_rate)
    nco = nco_for(plan)
    cap = 1 << 13
    mask = cap - 1
    stream = busy_stream(mode, rate, sum(LENGTHS), seed % (1 << 32))
    cpu_ring = np.zeros(2 * cap, dtype=np.float32)
    cpu_state = np.zeros(8, dtype=np.uint32)
    gpu_ring = torch.zeros(2 * cap, dtype=torch.float32, device="cuda")
    gpu_state = torch.zeros(8, dtype=torch.int32, device="cuda")
    nco_d = to_dev(nco.view(np.float32))
    at, taken = cap - 700, 0
    for d, length in enumerate(LENGTHS):
        slots = (at + np.arange(length)) & mask
        piece = stream[taken:taken + length]
        cpu_ring.view(np.complex64)[slots] = piece
        gpu_ring.view(torch.complex64)[torch.from_numpy(slots).cuda()] = \
            torch.from_numpy(piece).cuda()
        shift = float(np.float32(-TWO_PI * 120.0 / rate)) if d == 6 else 0.0
        p = plan.carrier_block(mask, at, length, reset=(d == 4), shift=shift)
        print(f"carrier mode {mode} dispatch {d} of {length} at {at}", flush=True)
        oracle.vrx_carrier(plan.plan.carrier, p, nco, cpu_ring, cpu_state)
        dev.vrx_carrier(plan.plan.carrier, p, nco_d, gpu_ring, gpu_state)
        assert_bit_exact(from_dev(gpu_ring), cpu_ring, what=f"ring after dispatch {d}")
        assert np.array_equal(from_dev(gpu_state).view(np.uint32), cpu_state), d
        at += length
        taken += length
    assert cpu_state[5] == 1 and cpu_state[0] != 0

    # and the detector after it (SAM's real part with AM's DC window, DSB's
    # product detector) on the rotated ring
    hist = plan.history
    first = cap - 700 + hist
    count = sum(LENGTHS) - hist
    dp = _lib.VrxDemodParams(mask, first & mask, count, plan.plan.demod_gain)
    want = oracle.vrx_demod(plan.demod, dp, cpu_ring, plan.weights)
    audio = torch.zeros(count * plan.demod.channels, dtype=torch.float32, device="cuda")
    dev.vrx_demod(plan.demod, dp, gpu_ring, to_dev(plan.weights), audio)
    assert_bit_exact(from_dev(audio), want, what=f"demod mode {mode} after the loop")


@pytest.mark.parametrize("log2", [2, 5, 16, 18])
def test_carrier_on_any_table_size(dev, log2):
    """The table's first quarter in LDS (log2 2..16) and the table in memory
    (above 16) give the twin's bits."""
    from revenant_amd.dsp import build_nco_table
    plan = sam_plan()
    rate = float(plan.plan.demod_rate)
    nco = build_nco_table(log2)
    cfg = _lib.VrxCarrierConfig(plan.plan.carrier.power, log2)
    cap = 1 << 12
    stream = busy_stream(DEMOD_SAM, rate, 3000, 5 + log2)
    cpu_ring = np.zeros(2 * cap, dtype=np.float32)
    cpu_ring[:6000] = stream.view(np.float32)
    gpu_ring = to_dev(cpu_ring)
    cpu_state = np.zeros(8, dtype=np.uint32)
    gpu_state = torch.zeros(8, dtype=torch.int32, device="cuda")
    nco_d = to_dev(nco.view(np.float32))
    for at, length in ((0, 1000), (1000, 2000)):
        p = plan.carrier_block(cap - 1, at, length)
        print(f"carrier table 2^{log2} at {at} x {length}", flush=True)
        oracle.vrx_carrier(cfg, p, nco, cpu_ring, cpu_state)
        dev.vrx_carrier(cfg, p, nco_d, gpu_ring, gpu_state)
        assert_bit_exact(from_dev(gpu_ring), cpu_ring, what=f"table 2^{log2}")
        assert np.array_equal(from_dev(gpu_state).view(np.uint32), cpu_state)


def _items_tensor(items):
    raw = b"".join(bytes(it) for it in items)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).cuda()


def test_batched_loops_match_each_receiver_alone(dev):
    g = np.random.default_rng(0x43415252)
    plans = {DEMOD_SAM: sam_plan(), DEMOD_DSB: dsb_plan()}
    nco = nco_for(plans[DEMOD_SAM])
    nco_d = to_dev(nco.view(np.float32))
    n = 130  # two full waves and a partial one
    rx = []
    for k in range(n):
        mode = DEMOD_SAM if k % 2 == 0 else DEMOD_DSB
        cap = 1 << int(g.integers(11, 14))
        rate = float(plans[mode].plan.demod_rate)
        stream = busy_stream(mode, rate, 3 * cap, int(g.integers(1, 1 << 30)))
        rx.append(dict(mode=mode, cap=cap, stream=stream, at=int(g.integers(0, cap)), taken=0,
                       cpu_ring=np.zeros(2 * cap, dtype=np.float32),
                       cpu_state=np.zeros(8, dtype=np.uint32),
                       gpu_ring=torch.zeros(2 * cap, dtype=torch.float32, device="cuda"),
                       gpu_state=torch.zeros(8, dtype=torch.int32, device="cuda")))
    for launch in range(3):
        items = []
        for k, r in enumerate(rx):
            plan = plans[r["mode"]]
            mask = r["cap"] - 1
            length = 0 if (k + launch) % 17 == 0 else int(g.integers(1, r["cap"] // 2))
            slots = (r["at"] + np.arange(length)) & mask
            piece = r["stream"][r["taken"]:r["taken"] + length]
            r["cpu_ring"].view(np.complex64)[slots] = piece
            if length:
                r["gpu_ring"].view(torch.complex64)[torch.from_numpy(slots).cuda()] = \
                    torch.from_numpy(piece).cuda()
            reset = launch == 2 and k % 5 == 0
            p = plan.carrier_block(mask, r["at"], length, reset=reset)
            oracle.vrx_carrier(plan.plan.carrier, p, nco, r["cpu_ring"], r["cpu_state"])
            it = _lib.VrxCarrierItem(plan.plan.carrier, p, r["gpu_ring"].data_ptr(), r["cap"],
                                     r["gpu_state"].data_ptr(), 0)
            _lib.check(_lib.lib().rvx_vrx_carrier_item_check(C.byref(it), nco.size))
            items.append(it)
            r["at"] += length
            r["taken"] += length
        dev_items = _items_tensor(items)
        print(f"carrier batch launch {launch}: {n} receivers", flush=True)
        dev.vrx_carrier_batch(dev_items, n, nco_d)
        torch.cuda.synchronize()
        for k, r in enumerate(rx):
            assert_bit_exact(from_dev(r["gpu_ring"]), r["cpu_ring"],
                             what=f"receiver {k} ring, launch {launch}")
            assert np.array_equal(from_dev(r["gpu_state"]).view(np.uint32), r["cpu_state"]), \
                (k, launch)
    assert sum(int(r["cpu_state"][5]) for r in rx) > n // 2  # most loops locked


def test_item_check_refuses_mismatched_sizes(dev):
    plan = sam_plan()
    nco = nco_for(plan)
    p = plan.carrier_block(1023, 0, 100)
    ok = _lib.VrxCarrierItem(plan.plan.carrier, p, 2, 1024, 3, 0)
    assert _lib.lib().rvx_vrx_carrier_item_check(C.byref(ok), nco.size) == 0
    assert _lib.lib().rvx_vrx_carrier_item_check(C.byref(ok), nco.size // 2) != 0
    for field, value in (("fine_samples", 1000), ("state", None), ("fine_ring", None)):
        bad = _lib.VrxCarrierItem.from_buffer_copy(bytes(ok))
        setattr(bad, field, value)
        assert _lib.lib().rvx_vrx_carrier_item_check(C.byref(bad), nco.size) != 0, field
    ring = torch.zeros(2 * 512, dtype=torch.float32, device="cuda")
    state = torch.zeros(8, dtype=torch.int32, device="cuda")
    with pytest.raises(_lib.RvxError):  # a ring smaller than the mask implies
        dev.vrx_carrier(plan.plan.carrier, p, to_dev(nco.view(np.float32)), ring, state)
