"""GPU vs oracle, bit for bit: the impulse blanker's detect and apply passes
across a ring wrap, after tests/reference/test_noise.cpp:163-268, over counts
that leave partial and single-lane workgroups and over the shortest and
longest windows and a long guard."""
import numpy as np
import pytest

from helpers import assert_bit_exact, from_dev, to_dev
from oracle import oracle
from revenant_amd import _lib
from revenant_amd.dsp import BLANK_APPLY, BLANK_DETECT
from tests.blanker_cases import (OUT_RING, apply_params, config_144k, impulsive_channels,
                                 wrapping_params)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0x4E4F495345000001


def _with(c, **fields):
    out = _lib.BlankerConfig.from_buffer_copy(bytes(c))
    for k, v in fields.items():
        setattr(out, k, v)
    return out


@pytest.mark.parametrize("count,shape", [(1500, {}), (257, {}), (1, {}),
                                         (1500, {"window": 16}), (1500, {"window": 256}),
                                         (1200, {"guard": 300, "hang": 5, "lead": 9})])
def test_detect_and_apply_match_twins_through_a_wrap(dev, count, shape):
    c = _with(config_144k(), **shape)
    ch = impulsive_channels(SEED + count)
    p, first = wrapping_params(count)
    d_ch = to_dev(ch.reshape(-1))
    want = np.zeros(OUT_RING, np.float32)
    oracle.blank_detect(_with(c, pass_=BLANK_DETECT), p, ch, want)
    got = torch.zeros(OUT_RING, dtype=torch.float32, device="cuda")
    print(f"blank detect count {count} {shape}", flush=True)
    dev.blank_detect(c, p, d_ch, got)
    assert_bit_exact(from_dev(got), want, SEED, "blanker flags")
    if count == 1500 and not shape:
        assert want.sum() > 20
    if count <= c.hang + c.lead:
        return
    a = apply_params(p, first, c)
    w_blanked = np.full((OUT_RING, 2), 3.0, np.float32)
    oracle.blank_apply(_with(c, pass_=BLANK_APPLY), a, ch, want, w_blanked)
    g_blanked = torch.full((2 * OUT_RING,), 3.0, dtype=torch.float32, device="cuda")
    print(f"blank apply count {a.count}", flush=True)
    dev.blank_apply(c, a, d_ch, to_dev(want), g_blanked)
    assert_bit_exact(from_dev(g_blanked), w_blanked.reshape(-1), SEED, "blanked ring")


def test_refusals_launch_nothing(dev):
    c = config_144k()
    p, _ = wrapping_params()
    ch = to_dev(impulsive_channels(SEED).reshape(-1))
    short = torch.zeros(OUT_RING - 1, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.RvxError):
        dev.blank_detect(c, p, ch, short)
    bad = _lib.BlankerParams.from_buffer_copy(bytes(p))
    bad.chan_base = 3 * 4096  # base + ring runs past the channel buffer
    with pytest.raises(_lib.RvxError):
        dev.blank_detect(c, bad, ch, torch.zeros(OUT_RING, dtype=torch.float32, device="cuda"))
