"""GPU vs oracle, bit for bit: convert, branch filter, transform, fused chain.

Mirrors tests/reference/test_convert.cpp, test_pfb.cpp (227-513) and
test_cmul.cpp of the reference: seeded uniform input, adversarial values,
several grids, and the device denormal probe (gpu_fixture.h:117-136).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import (SEED_BRANCH, SEED_CHAIN, SEED_FFT, adversarial, assert_bit_exact, from_dev,
                     rng, to_dev, uniform_complex)
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (FORMAT_CS16, FORMAT_CU8, GridParams, build_twiddles, prescale_taps,
                              design_prototype, ema_frame_decay)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_device_flushes_denormals(dev):
    """cmul with a product of 1e-40 must give exact zero (gpu_fixture.h:117-136)."""
    a = np.array([1e-20, 1e-20, 3e-20, -2e-20], dtype=np.float32)
    b = np.array([1e-20, 0.0, 1e-20, 1e-20], dtype=np.float32)
    out = torch.empty(4, dtype=torch.float32, device="cuda")
    dev.cmul(to_dev(a), to_dev(b), out)
    got = from_dev(out)
    assert np.all(got == 0.0), got
    assert_bit_exact(got, oracle.cmul(a, b), what="cmul denormal probe")


def test_cmul_matches_twin(dev):
    a = adversarial(0x434D554C00000001, 4096)
    b = uniform_complex(0x434D554C00000002, 4096, 3.0)
    out = torch.empty(a.size, dtype=torch.float32, device="cuda")
    dev.cmul(to_dev(a), to_dev(b), out)
    assert_bit_exact(from_dev(out), oracle.cmul(a, b), what="cmul")


@pytest.mark.parametrize("fmt,bps", [(0, 2), (1, 2), (2, 4), (3, 6)])
@pytest.mark.parametrize("src_offset,count", [(0, 1000), (3, 4093), (17, 1)])
def test_convert_matches_twin(dev, fmt, bps, src_offset, count):
    g = rng(0x434F4E5600000000 + fmt)
    nbytes = ((src_offset + count) * bps + 3) // 4 * 4
    packed = g.integers(0, 256, size=nbytes, dtype=np.uint8)
    cap = 4096
    p = _lib.ConvertParams(cap - 1, 0xFFFFFF00, src_offset, count)  # wraps the ring
    ring_cpu = np.full(2 * cap, 7.0, dtype=np.float32)
    oracle.convert(fmt, packed, p, ring_cpu)
    ring = to_dev(np.full(2 * cap, 7.0, dtype=np.float32))
    dev.convert(fmt, to_dev(packed), p, ring)
    assert_bit_exact(from_dev(ring), ring_cpu, what=f"convert fmt {fmt}")


def test_convert_every_cu8_code(dev):
    codes = np.arange(256, dtype=np.uint8)
    packed = np.stack([codes, codes[::-1]], axis=1).ravel()
    p = _lib.ConvertParams(255, 0, 0, 256)
    ring = torch.zeros(512, dtype=torch.float32, device="cuda")
    dev.convert(FORMAT_CU8, to_dev(packed), p, ring)
    cpu = np.zeros(512, dtype=np.float32)
    oracle.convert(FORMAT_CU8, packed, p, cpu)
    assert_bit_exact(from_dev(ring), cpu, what="cu8 codes")


def test_convert_refuses_short_source(dev):
    p = _lib.ConvertParams(1023, 0, 0, 3)
    ring = torch.zeros(2048, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.RvxError):
        dev.convert(FORMAT_CU8, to_dev(np.zeros(6, dtype=np.uint8)), p, ring)


def _ring_for(grid: GridParams, blocks: int, seed: int, adv: bool):
    need = grid.prototype_length + (blocks - 1) * grid.decimation
    cap = 1 << (need - 1).bit_length()
    data = adversarial(seed, cap) if adv else uniform_complex(seed, cap)
    return cap, data


GRIDS = [GridParams(8, 4, 8), GridParams(8, 17, 4), GridParams(64, 17, 32),
         GridParams(64, 16, 64), GridParams(256, 16, 128), GridParams(4096, 16, 2048),
         GridParams(2048, 17, 1024)]


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: f"M{g.channels}L{g.taps_per_branch}D{g.decimation}")
@pytest.mark.parametrize("adv", [False, True], ids=["uniform", "adversarial"])
def test_branch_matches_twin(dev, grid, adv):
    blocks = 24
    cap, ring = _ring_for(grid, blocks, SEED_BRANCH, adv)
    taps = design_prototype(grid).taps
    p = _lib.PfbBranchParams(cap - 1, 0xFFFFFFF0 if grid.channels < 1024 else 12345, blocks)
    cpu = oracle.pfb_branch(grid.c(), taps, ring, p)
    out = torch.empty(2 * blocks * grid.channels, dtype=torch.float32, device="cuda")
    dev.pfb_branch(grid, to_dev(taps), to_dev(ring), p, out)
    assert_bit_exact(from_dev(out), cpu, SEED_BRANCH, f"branch {grid}")


@pytest.mark.parametrize("M,D", [(2, 2), (4, 2), (8, 8), (16, 8), (64, 32), (64, 64),
                                 (512, 256), (4096, 2048), (4096, 4096), (16384, 8192)])
@pytest.mark.parametrize("adv", [False, True], ids=["uniform", "adversarial"])
def test_fft_matches_twin(dev, M, D, adv):
    blocks = 13
    tw = build_twiddles(M)
    branch = adversarial(SEED_FFT, blocks * M) if adv else uniform_complex(SEED_FFT, blocks * M)
    ring_blocks = 16
    p = _lib.PfbFftParams(M, D, M.bit_length() - 1, 0xFFFFFFFB, blocks, ring_blocks,
                          ring_blocks - 1)
    cpu = np.full(2 * M * ring_blocks, 5.0, dtype=np.float32)
    oracle.pfb_fft(p, branch, tw, cpu)
    out = to_dev(np.full(2 * M * ring_blocks, 5.0, dtype=np.float32))
    dev.pfb_fft(p, to_dev(branch), to_dev(tw.view(np.float32)), out)
    assert_bit_exact(from_dev(out), cpu, SEED_FFT, f"pfb_fft M={M} D={D}")


@pytest.mark.parametrize("size", [2, 4, 8, 16, 32, 64, 128, 1024, 2048, 4096, 8192, 16384])
def test_fft_radix2_matches_twin(dev, size):
    batch = 5
    tw = build_twiddles(size)
    x = uniform_complex(SEED_FFT + size, batch * size)
    cpu = oracle.fft_radix2(tw, x, size, batch)
    out = torch.empty(2 * batch * size, dtype=torch.float32, device="cuda")
    dev.fft_radix2(to_dev(tw.view(np.float32)), to_dev(x), out, size, batch)
    assert_bit_exact(from_dev(out), cpu, SEED_FFT, f"fft_radix2 {size}")


def _codes(fmt: int, n: int, seed: int) -> np.ndarray:
    g = rng(seed)
    if fmt == FORMAT_CU8:
        return g.integers(0, 256, size=2 * n, dtype=np.uint8)
    return g.integers(-32768, 32768, size=2 * n, dtype=np.int16).view(np.uint8)


CHAINS = [
    # (format, grid, blocks, frame) -- the C1 grid, the C2 grid, and odd shapes
    (FORMAT_CU8, GridParams(64, 16, 64), 256, 16),
    (FORMAT_CU8, GridParams(64, 16, 32), 96, 24),
    (FORMAT_CS16, GridParams(4096, 16, 2048), 64, 16),
    (FORMAT_CS16, GridParams(4096, 17, 2048), 32, 8),
    (FORMAT_CS16, GridParams(1024, 16, 512), 48, 0),
    (FORMAT_CS16, GridParams(512, 16, 256), 64, 4),
    (FORMAT_CS16, GridParams(8, 5, 4), 40, 10),
    (FORMAT_CU8, GridParams(256, 12, 256), 32, 4),
]


@pytest.mark.parametrize("fmt,grid,blocks,frame", CHAINS,
                         ids=lambda v: str(v) if not isinstance(v, GridParams) else
                         f"M{v.channels}L{v.taps_per_branch}D{v.decimation}")
def test_chain_matches_composed_twins(dev, fmt, grid, blocks, frame):
    """The fused launch equals convert -> branch -> fft -> ema_partial."""
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + 3
    samples = (first_block + blocks) * D + 7
    packed = _codes(fmt, samples, SEED_CHAIN + M)
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks - 1).bit_length()
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(fmt, M, grid.taps_per_branch, D, first_block, blocks, ring_blocks, frame,
                       alpha, ema_frame_decay(alpha, frame) if frame else 0.0)
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    desc = _lib.ChainDesc(fmt, M, grid.taps_per_branch, D, 0, samples, first_block, blocks,
                          ring_blocks, frame, alpha)
    ring = torch.full((2 * M * ring_blocks,), 3.0, dtype=torch.float32, device="cuda")
    partial = (torch.zeros((blocks // frame) * M, dtype=torch.float32, device="cuda")
               if frame else None)
    dev.chain(desc, to_dev(packed), to_dev(taps), to_dev(tw.view(np.float32)), ring, partial)
    got = from_dev(ring)
    # Slots the launch does not write keep the fill in both (oracle starts at 0).
    written = np.zeros((M, ring_blocks), dtype=bool)
    for b in range(blocks):
        written[:, (first_block + b) & (ring_blocks - 1)] = True
    mask = np.repeat(written.ravel(), 2)
    assert_bit_exact(got[mask], ring_cpu[mask], SEED_CHAIN, f"chain ring {grid}")
    if frame:
        assert_bit_exact(from_dev(partial), partial_cpu, SEED_CHAIN, f"chain ema {grid}")


# The two-times-oversampled grids (D = M/2) with 16 or 17 taps and no frame or
# a four-block frame run on the pair-streaming kernel: every transform shape
# it instantiates, tails of 1..3 blocks, chunks that do not start on a
# four-block ring boundary, cu8 and cs16.
PAIR_CASES = [
    (FORMAT_CU8, GridParams(4, 16, 2), 37, 0, 2),
    (FORMAT_CS16, GridParams(8, 17, 4), 40, 4, 1),
    (FORMAT_CU8, GridParams(16, 16, 8), 41, 0, 3),
    (FORMAT_CS16, GridParams(64, 16, 32), 64, 4, 0),
    (FORMAT_CU8, GridParams(64, 17, 32), 30, 0, 1),
    (FORMAT_CS16, GridParams(256, 16, 128), 35, 0, 0),
    (FORMAT_CS16, GridParams(512, 16, 256), 70, 0, 2),
    (FORMAT_CU8, GridParams(512, 17, 256), 16, 4, 0),
    (FORMAT_CS16, GridParams(1024, 16, 512), 22, 0, 3),
    (FORMAT_CS16, GridParams(2048, 17, 1024), 12, 4, 0),
    (FORMAT_CS16, GridParams(4096, 16, 2048), 36, 4, 0),
    (FORMAT_CU8, GridParams(4096, 16, 2048), 9, 0, 1),
    (FORMAT_CS16, GridParams(4096, 17, 2048), 13, 0, 2),
]


@pytest.mark.parametrize("fmt,grid,blocks,frame,skew", PAIR_CASES,
                         ids=[f"{'cu8' if c[0] == FORMAT_CU8 else 'cs16'}-M{c[1].channels}"
                              f"L{c[1].taps_per_branch}-{c[2]}b-f{c[3]}-s{c[4]}"
                              for c in PAIR_CASES])
def test_pair_chain_matches_composed_twins(dev, fmt, grid, blocks, frame, skew):
    """The pair-streaming launch equals convert -> branch -> fft -> ema_partial,
    also with prescaled cs16 taps."""
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + skew
    samples = (first_block + blocks) * D + 5
    packed = _codes(fmt, samples, SEED_CHAIN + 7 * M + skew)
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks + 3).bit_length()
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(fmt, M, grid.taps_per_branch, D, first_block, blocks, ring_blocks, frame,
                       alpha, ema_frame_decay(alpha, frame) if frame else 0.0)
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    written = np.zeros((M, ring_blocks), dtype=bool)
    for b in range(blocks):
        written[:, (first_block + b) & (ring_blocks - 1)] = True
    mask = np.repeat(written.ravel(), 2)
    scaled_cases = (False, True) if fmt == FORMAT_CS16 else (False,)
    for scaled in scaled_cases:
        desc = _lib.ChainDesc(fmt, M, grid.taps_per_branch, D, 0, samples, first_block, blocks,
                              ring_blocks, frame, alpha)
        desc.taps_scaled = int(scaled)
        use = prescale_taps(taps) if scaled else taps
        ring = torch.full((2 * M * ring_blocks,), 3.0, dtype=torch.float32, device="cuda")
        partial = (torch.zeros((blocks // frame) * M, dtype=torch.float32, device="cuda")
                   if frame else None)
        print(f"pair chain M={M} taps={grid.taps_per_branch} blocks={blocks} frame={frame} "
              f"skew={skew} scaled={scaled}", flush=True)
        dev.chain(desc, to_dev(packed), to_dev(use), to_dev(tw.view(np.float32)), ring, partial)
        got = from_dev(ring)
        assert_bit_exact(got[mask], ring_cpu[mask], SEED_CHAIN, f"pair chain ring {grid}")
        assert np.all(got[~mask] == 3.0), "a slot outside the launch's blocks was written"
        if frame:
            assert_bit_exact(from_dev(partial), partial_cpu, SEED_CHAIN, f"pair chain ema {grid}")


def test_chain_refuses_missing_history(dev):
    grid = GridParams(64, 16, 64)
    desc = _lib.ChainDesc(FORMAT_CU8, 64, 16, 64, 0, 10000, 2, 8, 8, 0, 0.0625)
    ring = torch.zeros(2 * 64 * 8, dtype=torch.float32, device="cuda")
    taps = to_dev(design_prototype(grid).taps)
    tw = to_dev(build_twiddles(64).view(np.float32))
    with pytest.raises(_lib.RvxError):
        dev.chain(desc, to_dev(np.zeros(20000, dtype=np.uint8)), taps, tw, ring)


@pytest.mark.parametrize("grid,blocks", [(GridParams(512, 16, 256), 64),
                                         (GridParams(4096, 16, 2048), 24)],
                         ids=["M512", "M4096"])
def test_chain_stores_selected_channels(dev, grid, blocks):
    """With a channel-row table the launch stores only the chosen channels,
    each in its row of a compact ring, bit-identical to the full ring's."""
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + 5
    samples = (first_block + blocks) * D + 3
    packed = _codes(FORMAT_CS16, samples, SEED_CHAIN + 7 * M)
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks - 1).bit_length()
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(FORMAT_CS16, M, 16, D, first_block, blocks, ring_blocks, 4, alpha,
                       ema_frame_decay(alpha, 4))
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    chosen = [M - 1, 3, 0, M // 2 + 1, 77]
    rows = np.full(M, 0xFFFFFFFF, dtype=np.uint32)
    rows[9] = len(chosen) + 4   # beyond ring_rows: not stored either
    for r, k in enumerate(chosen):
        rows[k] = r
    desc = _lib.ChainDesc(FORMAT_CS16, M, 16, D, 0, samples, first_block, blocks, ring_blocks,
                          4, alpha)
    rows_dev = to_dev(rows.view(np.int32))
    desc.ring_rows = len(chosen)
    desc.channel_rows = rows_dev.data_ptr()
    ring = torch.full((2 * len(chosen) * ring_blocks,), 3.0, dtype=torch.float32, device="cuda")
    partial = torch.zeros((blocks // 4) * M, dtype=torch.float32, device="cuda")
    dev.chain(desc, to_dev(packed), to_dev(taps), to_dev(tw.view(np.float32)), ring, partial)
    got = from_dev(ring).reshape(len(chosen), ring_blocks, 2)
    full = ring_cpu.reshape(M, ring_blocks, 2)
    slots = [(first_block + b) & (ring_blocks - 1) for b in range(blocks)]
    for r, k in enumerate(chosen):
        assert_bit_exact(got[r, slots], full[k, slots], SEED_CHAIN, f"row {r} = channel {k}")
    assert_bit_exact(from_dev(partial), partial_cpu, SEED_CHAIN, "spectrum over every channel")



# (format, grid, blocks, frame, selected rows, padding): the pair-streaming
# kernel with prescaled taps, the general kernel, and a selected-row ring.
PITCH_CASES = [
    (FORMAT_CS16, GridParams(4096, 16, 2048), 24, 4, None, 32),
    (FORMAT_CU8, GridParams(64, 16, 64), 40, 4, None, 4),
    (FORMAT_CS16, GridParams(512, 16, 256), 48, 0, [511, 3, 0, 257, 77], 32),
]


@pytest.mark.parametrize("fmt,grid,blocks,frame,chosen,pad", PITCH_CASES,
                         ids=["M4096-pair-pad32", "M64-general-pad4", "M512-rows-pad32"])
def test_chain_ring_pitch_keeps_rows_bit_exact(dev, fmt, grid, blocks, frame, chosen, pad):
    """Rows ring_blocks + pad apart: every row holds the twins' channel bits,
    and the padding between rows is never written."""
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + 3
    samples = (first_block + blocks) * D + 5
    packed = _codes(fmt, samples, SEED_CHAIN + 11 * M)
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks - 1).bit_length()
    pitch = ring_blocks + pad
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(fmt, M, grid.taps_per_branch, D, first_block, blocks, ring_blocks, frame,
                       alpha, ema_frame_decay(alpha, frame) if frame else 0.0)
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    full = ring_cpu.reshape(M, ring_blocks, 2)
    desc = _lib.ChainDesc(fmt, M, grid.taps_per_branch, D, 0, samples, first_block, blocks,
                          ring_blocks, frame, alpha)
    desc.ring_pitch = pitch
    use = taps
    if fmt == FORMAT_CS16 and chosen is None:
        desc.taps_scaled = 1
        use = prescale_taps(taps)
    rows_dev = None
    n_rows = M
    if chosen is not None:
        rows = np.full(M, 0xFFFFFFFF, dtype=np.uint32)
        for r, k in enumerate(chosen):
            rows[k] = r
        rows_dev = to_dev(rows.view(np.int32))
        desc.ring_rows = len(chosen)
        desc.channel_rows = rows_dev.data_ptr()
        n_rows = len(chosen)
    ring = torch.full((2 * n_rows * pitch,), 3.0, dtype=torch.float32, device="cuda")
    partial = (torch.zeros((blocks // frame) * M, dtype=torch.float32, device="cuda")
               if frame else None)
    print(f"ring pitch M={M} rows={n_rows} ring_blocks={ring_blocks} pitch={pitch}", flush=True)
    dev.chain(desc, to_dev(packed), to_dev(use), to_dev(tw.view(np.float32)), ring, partial)
    got = from_dev(ring).reshape(n_rows, pitch, 2)
    slots = [(first_block + b) & (ring_blocks - 1) for b in range(blocks)]
    channels = chosen if chosen is not None else list(range(M))
    for r, k in enumerate(channels):
        assert_bit_exact(got[r, slots], full[k, slots], SEED_CHAIN, f"row {r} = channel {k}")
    assert np.all(got[:, ring_blocks:] == 3.0), "the padding between rows was written"
    if frame:
        assert_bit_exact(from_dev(partial), partial_cpu, SEED_CHAIN, "spectrum with a pitch")
    # a pitch below the row length or off a multiple of 4 is refused
    for bad in (ring_blocks - 4, ring_blocks + 2):
        desc.ring_pitch = bad
        with pytest.raises(_lib.RvxError, match="pitch"):
            dev.chain(desc, to_dev(packed), to_dev(use), to_dev(tw.view(np.float32)), ring,
                      partial)


@pytest.mark.parametrize("grid,blocks", [(GridParams(4096, 16, 2048), 32),
                                         (GridParams(512, 16, 256), 64),
                                         (GridParams(1024, 17, 512), 24)],
                         ids=["M4096", "M512", "M1024L17"])
def test_chain_prescaled_taps_bit_exact(dev, grid, blocks):
    """cs16 against taps carrying the conversion's 2^-15: the same bits as the
    twins with the plain taps (a product of the same real numbers)."""
    from revenant_amd.dsp import prescale_taps

    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first_block = (N + D - 1) // D + 1
    samples = (first_block + blocks) * D + 5
    packed = _codes(FORMAT_CS16, samples, SEED_CHAIN + 3 * M)
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks - 1).bit_length()
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(FORMAT_CS16, M, grid.taps_per_branch, D, first_block, blocks, ring_blocks,
                       8, alpha, ema_frame_decay(alpha, 8))
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    desc = _lib.ChainDesc(FORMAT_CS16, M, grid.taps_per_branch, D, 0, samples, first_block,
                          blocks, ring_blocks, 8, alpha)
    desc.taps_scaled = 1
    ring = torch.full((2 * M * ring_blocks,), 3.0, dtype=torch.float32, device="cuda")
    partial = torch.zeros((blocks // 8) * M, dtype=torch.float32, device="cuda")
    dev.chain(desc, to_dev(packed), to_dev(prescale_taps(taps)), to_dev(tw.view(np.float32)),
              ring, partial)
    got = from_dev(ring).reshape(M, ring_blocks, 2)
    slots = [(first_block + b) & (ring_blocks - 1) for b in range(blocks)]
    assert_bit_exact(got[:, slots], ring_cpu.reshape(M, ring_blocks, 2)[:, slots], SEED_CHAIN,
                     f"prescaled ring {grid}")
    assert_bit_exact(from_dev(partial), partial_cpu, SEED_CHAIN, f"prescaled ema {grid}")
