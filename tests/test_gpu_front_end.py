"""GPU vs oracle, bit for bit: front-end correction.

After tests/reference/test_iq_correction.cpp:91-164 (the moments and the
correction kernels against their twins, signal and adversarial rings, through
a wrap, real and identity coefficients) and test_front_end_correction.cpp
(the estimator's schedule, here with the device doing the per-sample work),
plus this build's two fusions: the moments read straight from native codes,
and the correction folded into the channelizer's conversion (the chain with a
correction equals convert -> iq_correct -> branch filter -> transform, with
the coefficients switching at a stated sample).
"""
import numpy as np
import pytest

from helpers import adversarial, assert_bit_exact, from_dev, rng, to_dev, uniform_complex
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (FORMAT_CS16, FORMAT_CU8, GridParams, build_twiddles,
                              design_prototype)
from revenant_amd.front_end import FrontEndCorrector

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0x4951434F52520001
CAP = 1 << 16
OFFSET = CAP - 5000           # blocks wrap the ring
COUNT = 40_000 + 37           # ends in a partial chunk
REAL = (0.0213, -0.0097, -0.0524078, 0.9536746)


def _ring(adv: bool) -> np.ndarray:
    if adv:
        return adversarial(SEED + 1, CAP)
    t = np.arange(CAP)
    z = 0.3 * np.exp(2j * np.pi * 0.013 * t) + 0.01 + 0.005j
    z = z + (rng(SEED).standard_normal(CAP) + 1j * rng(SEED + 7).standard_normal(CAP)) * 0.01
    return z.astype(np.complex64).view(np.float32)


@pytest.mark.parametrize("adv", [False, True], ids=["signal", "adversarial"])
def test_moments_match_twin(dev, adv):
    ring = _ring(adv)
    want = oracle.iq_moments(CAP - 1, OFFSET, COUNT, ring).ravel()
    sums = torch.full((want.size + 10,), 7.0, dtype=torch.float32, device="cuda")
    dev.iq_moments(_lib.IqMomentsParams(CAP - 1, OFFSET, COUNT), to_dev(ring), sums)
    got = from_dev(sums)
    assert_bit_exact(got[:want.size], want, SEED, "iq_moments")
    assert np.all(got[want.size:] == 7.0)  # nothing past the chunks is written


@pytest.mark.parametrize("adv", [False, True], ids=["signal", "adversarial"])
@pytest.mark.parametrize("coeffs", [REAL, (0.0, 0.0, 0.0, 1.0)], ids=["real", "identity"])
@pytest.mark.parametrize("in_place", [True, False], ids=["in-place", "two-rings"])
def test_correct_matches_twin(dev, adv, coeffs, in_place):
    ring = _ring(adv)
    want = oracle.iq_correct(CAP - 1, OFFSET, COUNT, *coeffs, ring)
    p = _lib.IqCorrectParams(CAP - 1, OFFSET, COUNT, *coeffs)
    src = to_dev(ring)
    if in_place:
        dev.iq_correct(p, src)
        got = from_dev(src)
    else:
        out = to_dev(ring)  # slots outside the block keep the input in both
        dev.iq_correct(p, src, out)
        got = from_dev(out)
    assert_bit_exact(got, want, SEED, f"iq_correct {coeffs}")


@pytest.mark.parametrize("fmt", [FORMAT_CS16, FORMAT_CU8], ids=["cs16", "cu8"])
def test_moments_over_codes_equal_moments_of_the_converted_ring(dev, fmt):
    g = rng(SEED + fmt)
    n = 70_000
    if fmt == FORMAT_CU8:
        packed = g.integers(0, 256, size=2 * n, dtype=np.uint8)
    else:
        packed = g.integers(-32768, 32768, size=2 * n, dtype=np.int16).view(np.uint8)
    ring = np.zeros(2 * (1 << 17), dtype=np.float32)
    oracle.convert(fmt, packed, _lib.ConvertParams((1 << 17) - 1, 0, 0, n), ring)
    first, count = 1234, 60_001
    want = oracle.iq_moments((1 << 17) - 1, first, count, ring).ravel()
    sums = torch.zeros(want.size, dtype=torch.float32, device="cuda")
    dev.iq_moments_codes(fmt, to_dev(packed), first, count, sums)
    assert_bit_exact(from_dev(sums), want, SEED, f"iq_moments_codes fmt {fmt}")


def _corrected_chain_oracle(fmt, grid, packed, samples, first_block, blocks, ring_blocks,
                            before, after, switch):
    """convert -> iq_correct (before up to `switch`, after from it) -> branch
    filter -> transform, on the oracle."""
    M, D = grid.channels, grid.decimation
    cap = 1 << (samples - 1).bit_length()
    ring = np.zeros(2 * cap, dtype=np.float32)
    words = np.zeros(-(-packed.size // 4) * 4, dtype=np.uint8)  # convert reads whole words
    words[:packed.size] = packed
    oracle.convert(fmt, words, _lib.ConvertParams(cap - 1, 0, 0, samples), ring)
    if switch > 0:
        ring = oracle.iq_correct(cap - 1, 0, switch, *before, ring)
    if switch < samples:
        ring = oracle.iq_correct(cap - 1, switch, samples - switch, *after, ring)
    taps = design_prototype(grid, beta=8.0).taps
    branch = oracle.pfb_branch(grid.c(), taps, ring,
                               _lib.PfbBranchParams(cap - 1, first_block * D, blocks))
    out = np.zeros(2 * M * ring_blocks, dtype=np.float32)
    oracle.pfb_fft(_lib.PfbFftParams(M, D, M.bit_length() - 1, first_block, blocks, ring_blocks,
                                     ring_blocks - 1), branch, build_twiddles(M), out)
    return out


@pytest.mark.parametrize("fmt,grid,blocks", [
    (FORMAT_CU8, GridParams(64, 16, 32), 96),
    (FORMAT_CS16, GridParams(512, 16, 256), 64),
    (FORMAT_CS16, GridParams(4096, 16, 2048), 40),
    (FORMAT_CS16, GridParams(1024, 13, 512), 24),
], ids=["cu8-M64", "cs16-M512", "cs16-M4096", "cs16-M1024L13"])
@pytest.mark.parametrize("where", ["inside", "before", "none"])
def test_chain_with_correction_matches_composed_twins(dev, fmt, grid, blocks, where):
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + 2
    samples = (first_block + blocks) * D + 5
    g = rng(SEED + M)
    if fmt == FORMAT_CU8:
        packed = g.integers(0, 256, size=2 * samples, dtype=np.uint8)
    else:
        packed = g.integers(-32768, 32768, size=2 * samples, dtype=np.int16).view(np.uint8)
    ring_blocks = 1 << (blocks - 1).bit_length()
    before = (0.011, -0.007, -0.03, 1.02)
    after = REAL
    # inside: the switch falls amid the history of the middle blocks
    switch = {"inside": (first_block + blocks // 2) * D - N // 3, "before": 0, "none": samples}[where]
    want = _corrected_chain_oracle(fmt, grid, packed, samples, first_block, blocks, ring_blocks,
                                   before, after, switch)
    desc = _lib.ChainDesc(fmt, M, grid.taps_per_branch, D, 0, samples, first_block, blocks,
                          ring_blocks, 0, 0.0)
    p_after = _lib.IqCorrectParams(0, 0, 0, *after)
    p_before = _lib.IqCorrectParams(0, 0, 0, *before)
    import ctypes as C
    desc.correction = C.cast(C.pointer(p_after), C.c_void_p)
    desc.correction_before = C.cast(C.pointer(p_before), C.c_void_p)
    desc.correction_from = switch
    ring = torch.zeros(2 * M * ring_blocks, dtype=torch.float32, device="cuda")
    taps = design_prototype(grid, beta=8.0).taps
    print(f"corrected chain fmt {fmt} M {M} blocks {blocks} switch {switch}", flush=True)
    dev.chain(desc, to_dev(packed), to_dev(taps), to_dev(build_twiddles(M).view(np.float32)),
              ring)
    got = from_dev(ring)
    written = np.zeros((M, ring_blocks), dtype=bool)
    for b in range(blocks):
        written[:, (first_block + b) & (ring_blocks - 1)] = True
    mask = np.repeat(written.ravel(), 2)
    assert_bit_exact(got[mask], want[mask], SEED, f"corrected chain {grid} switch {where}")


def test_chain_refuses_correction_with_prescaled_taps(dev):
    import ctypes as C
    grid = GridParams(64, 16, 32)
    desc = _lib.ChainDesc(FORMAT_CS16, 64, 16, 32, 0, 4096, 40, 8, 8, 0, 0.0)
    desc.taps_scaled = 1
    p = _lib.IqCorrectParams(0, 0, 0, *REAL)
    desc.correction = C.cast(C.pointer(p), C.c_void_p)
    ring = torch.zeros(2 * 64 * 8, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.RvxError):
        dev.chain(desc, torch.zeros(4096, dtype=torch.int32, device="cuda"),
                  to_dev(design_prototype(grid, beta=8.0).taps),
                  to_dev(build_twiddles(64).view(np.float32)), ring)


def test_estimator_schedule_on_the_device_matches_the_oracle(dev):
    """Moments and corrections on the device, the estimator between them on
    the host: the same corrected stream, bit for bit, as the oracle run."""
    rate, block, in_flight, cap = 2_400_000, 32768, 3, 1 << 16
    n = 24 * block
    t = np.arange(n, dtype=np.float64)
    ang = 2 * np.pi * 5000 * t / rate
    r = rng(SEED + 99)
    i_r = 0.1 * np.cos(ang) + 0.02 + 1e-4 * r.standard_normal(n)
    phi = 3.0 * np.pi / 180.0
    q_r = 1.05 * (0.1 * np.sin(ang) * np.cos(phi) + 0.1 * np.cos(ang) * np.sin(phi)) + 0.01 + \
        1e-4 * r.standard_normal(n)
    stream = np.stack([i_r, q_r], axis=1).astype(np.float32)

    def run(device: bool):
        ring_h = np.zeros(2 * cap, dtype=np.float32)
        ring_d = torch.zeros(2 * cap, dtype=torch.float32, device="cuda")
        fe = FrontEndCorrector()
        pending = [None] * in_flight
        out = np.empty_like(stream)
        off = 0
        for k, first in enumerate(range(0, n, block)):
            slot = k % in_flight
            if pending[slot] is not None:
                sums, count = pending[slot]
                fe.update_sums(from_dev(sums) if device else sums, count, rate)
                pending[slot] = None
            idx = (off + np.arange(block)) & (cap - 1)
            if device:
                ring_d.view(-1, 2)[torch.from_numpy(idx).cuda()] = \
                    torch.from_numpy(stream[first:first + block]).cuda()
                sums = torch.zeros(((block + 63) // 64) * 5, dtype=torch.float32, device="cuda")
                dev.iq_moments(_lib.IqMomentsParams(cap - 1, off, block), ring_d, sums)
                pending[slot] = (sums, block)
            else:
                ring_h.reshape(-1, 2)[idx] = stream[first:first + block]
                pending[slot] = (oracle.iq_moments(cap - 1, off, block, ring_h), block)
            if fe.estimate().measured:
                p = fe.correction(True, True)
                p.ring_mask, p.offset, p.count = cap - 1, off, block
                if device:
                    dev.iq_correct(p, ring_d)
                else:
                    ring_h = oracle.iq_correct(cap - 1, off, block, p.dc_i, p.dc_q, p.cross,
                                               p.scale, ring_h)
            out[first:first + block] = (from_dev(ring_d).reshape(-1, 2) if device
                                        else ring_h.reshape(-1, 2))[idx]
            off = (off + block) & (cap - 1)
        return out, fe.estimate()

    got, est_d = run(True)
    want, est_h = run(False)
    assert_bit_exact(got, want, SEED, "corrected stream")
    assert est_d == est_h and est_d.iq_plausible
