"""GPU: the fused chain reading converted samples (RVX_FORMAT_CF32), the
input the engine's conditioning hands it.

Converted cu8/cs16 codes through the cf32 form must equal the codes through
the native forms bit for bit (conversion is exact or correctly rounded, and
the chain adds the same products in the same order), on the pair-streaming
and the general kernels; and arbitrary floats, with the −0 excision writes,
must equal the composed branch and transform twins (pfb_branch_reference.cpp:
22, pfb_fft_reference.cpp:143).
"""
import numpy as np
import pytest

from helpers import adversarial, assert_bit_exact, from_dev, rng, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import (FORMAT_CF32, FORMAT_CS16, FORMAT_CU8, GridParams, build_twiddles,
                              design_prototype, ema_frame_decay)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0xC0F32000


def _converted(fmt, packed):
    """The reference's conversion in float32 arithmetic (convert.cpp:126-153,
    235): cu8 (c - 127.5f) * 0x3C008081, cs16 c * 2^-15."""
    if fmt == FORMAT_CU8:
        c = packed.astype(np.float32)
        scale = np.array([0x3C008081], dtype=np.uint32).view(np.float32)[0]
        return ((c - np.float32(127.5)) * scale).astype(np.float32)
    return (packed.view(np.int16).astype(np.float32) * np.float32(2.0 ** -15)).astype(np.float32)


CASES = [  # (format, grid, blocks, frame): pair kernel, general kernel
    (FORMAT_CS16, GridParams(4096, 16, 2048), 24, 4),
    (FORMAT_CU8, GridParams(512, 16, 256), 37, 0),
    (FORMAT_CS16, GridParams(64, 17, 32), 40, 4),
    (FORMAT_CU8, GridParams(64, 16, 64), 48, 16),
    (FORMAT_CS16, GridParams(256, 12, 256), 20, 4),
]


@pytest.mark.parametrize("fmt,grid,blocks,frame", CASES,
                         ids=[f"{'cu8' if c[0] == FORMAT_CU8 else 'cs16'}-M{c[1].channels}"
                              f"L{c[1].taps_per_branch}D{c[1].decimation}-f{c[3]}" for c in CASES])
def test_cf32_chain_equals_the_codes_chain(dev, fmt, grid, blocks, frame):
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    first = (N + D - 1) // D + 1
    samples = (first + blocks) * D + 3
    g = rng(SEED + M)
    packed = (g.integers(0, 256, size=2 * samples, dtype=np.uint8) if fmt == FORMAT_CU8 else
              g.integers(-32768, 32768, size=2 * samples, dtype=np.int16).view(np.uint8))
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    ring_blocks = 1 << (blocks - 1).bit_length()
    alpha = np.float32(1.0 / 16.0)
    od = oracle.OChain(fmt, M, grid.taps_per_branch, D, first, blocks, ring_blocks, frame, alpha,
                       ema_frame_decay(alpha, frame) if frame else 0.0)
    ring_cpu, partial_cpu = oracle.chain(od, packed, taps, tw, threads=4)
    desc = _lib.ChainDesc(FORMAT_CF32, M, grid.taps_per_branch, D, 0, samples, first, blocks,
                          ring_blocks, frame, alpha)
    ring = torch.full((2 * M * ring_blocks,), 3.0, dtype=torch.float32, device="cuda")
    partial = (torch.zeros((blocks // frame) * M, dtype=torch.float32, device="cuda")
               if frame else None)
    dev.chain(desc, to_dev(_converted(fmt, packed)), to_dev(taps), to_dev(tw.view(np.float32)),
              ring, partial)
    written = np.zeros((M, ring_blocks), dtype=bool)
    for b in range(blocks):
        written[:, (first + b) & (ring_blocks - 1)] = True
    mask = np.repeat(written.ravel(), 2)
    assert_bit_exact(from_dev(ring)[mask], ring_cpu[mask], SEED, f"cf32 chain ring {grid}")
    if frame:
        assert_bit_exact(from_dev(partial), partial_cpu, SEED, f"cf32 chain partials {grid}")


@pytest.mark.parametrize("grid", [GridParams(4096, 16, 2048), GridParams(64, 16, 64)],
                         ids=["pair-M4096", "general-M64D64"])
def test_cf32_chain_equals_the_composed_twins_on_arbitrary_floats(dev, grid):
    M, D, N = grid.channels, grid.decimation, grid.prototype_length
    blocks = 12
    first = (N + D - 1) // D + 2
    samples = (first + blocks) * D + 1
    flat = adversarial(SEED + 7, samples)  # float32 pairs
    flat[0::194] = -0.0  # excised samples are -0 in both parts
    flat[1::194] = -0.0
    cap = 1 << (samples - 1).bit_length()
    ring_in = np.zeros(2 * cap, dtype=np.float32)
    ring_in[:flat.size] = flat
    taps = design_prototype(grid, beta=8.0).taps
    tw = build_twiddles(M)
    bp = _lib.PfbBranchParams(cap - 1, first * D, blocks)
    branch = oracle.pfb_branch(grid.c(), taps, ring_in, bp)
    ring_blocks = 16
    fp = _lib.PfbFftParams(M, D, M.bit_length() - 1, first, blocks, ring_blocks, ring_blocks - 1)
    want = np.full(2 * M * ring_blocks, 3.0, dtype=np.float32)
    oracle.pfb_fft(fp, branch, tw, want)
    desc = _lib.ChainDesc(FORMAT_CF32, M, grid.taps_per_branch, D, 0, samples, first, blocks,
                          ring_blocks, 0, np.float32(0.0))
    ring = to_dev(np.full(2 * M * ring_blocks, 3.0, dtype=np.float32))
    dev.chain(desc, to_dev(flat), to_dev(taps), to_dev(tw.view(np.float32)), ring)
    assert_bit_exact(from_dev(ring), want, SEED, f"cf32 chain vs composed twins {grid}")


def test_cf32_chain_refuses_a_correction(dev):
    grid = GridParams(64, 16, 32)
    desc = _lib.ChainDesc(FORMAT_CF32, 64, 16, 32, 0, 4096, 40, 8, 8, 0, np.float32(0.0))
    corr = _lib.IqCorrectParams()
    corr.scale = 1.0
    import ctypes as C

    desc.correction = C.cast(C.pointer(corr), C.c_void_p)
    x = torch.zeros(2 * 4096, dtype=torch.float32, device="cuda")
    ring = torch.zeros(2 * 64 * 8, dtype=torch.float32, device="cuda")
    taps = to_dev(design_prototype(grid, beta=8.0).taps)
    with pytest.raises(_lib.RvxError):
        dev.chain(desc, x, taps, to_dev(build_twiddles(64).view(np.float32)), ring)
