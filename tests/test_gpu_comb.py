"""GPU vs oracle, bit for bit: comb project and subtract through a wrap,
after tests/reference/test_comb_kernels.cpp:130-210 (the full 1024-line comb,
one and sixteen sub-blocks, a count that leaves a short last chunk; subtract
in place and into a second ring, and with no lines)."""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, to_dev
from oracle import oracle
from revenant_amd.dsp import comb_table
from tests.comb_lines import MASK, RING, make_comb, noise, wrapping_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0x434F4D4250


def _lines_dev(lines):
    return torch.frombuffer(bytearray(bytes(lines)), dtype=torch.uint8).cuda()


@pytest.mark.parametrize("sub_blocks", [1, 16])
@pytest.mark.parametrize("count", [49152, 49104])
def test_project_matches_twin(dev, sub_blocks, count):
    t = comb_table()
    lines = make_comb(1024, SEED)
    raw = noise(RING, SEED + 1, 0.05)
    p = wrapping_params(1024, sub_blocks)
    p.count = count
    want = oracle.comb_project(p, raw, lines, t)
    got = torch.zeros(2 * want.size, dtype=torch.float32, device="cuda")
    print(f"comb project sub_blocks {sub_blocks} count {count}", flush=True)
    dev.comb_project(p, to_dev(raw.view(np.float32)), _lines_dev(lines), 1024,
                     to_dev(t.view(np.float32)), got)
    assert_bit_exact(from_dev(got), want.view(np.float32), SEED, "comb projections")


@pytest.mark.parametrize("n_lines", [1024, 40, 0])
@pytest.mark.parametrize("in_place", [True, False], ids=["in-place", "two-rings"])
def test_subtract_matches_twin(dev, n_lines, in_place):
    t = comb_table()
    lines = make_comb(n_lines, SEED + 2)
    raw = noise(RING, SEED + 3, 0.05)
    p = wrapping_params(n_lines, 1)
    want_clean, want_removed = oracle.comb_subtract(p, raw, lines if n_lines else None, t, RING)
    src = to_dev(raw.view(np.float32))
    clean = src if in_place else to_dev(raw.view(np.float32))
    removed = torch.zeros(2 * RING, dtype=torch.float32, device="cuda")
    print(f"comb subtract lines {n_lines} in_place {in_place}", flush=True)
    dev.comb_subtract(p, src, _lines_dev(lines), n_lines, to_dev(t.view(np.float32)), clean,
                      removed)
    assert_bit_exact(from_dev(clean), want_clean, SEED, "cleaned")
    slots = (p.first + np.arange(p.count)) & MASK
    sel = np.stack([2 * slots, 2 * slots + 1], 1).ravel()
    assert_bit_exact(from_dev(removed)[sel], want_removed[sel], SEED, "removed")
