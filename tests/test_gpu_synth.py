"""The device scene render against its FP64 CPU evaluation (spot checks at the
start of a stream and 2^27 samples in, where exact phase reduction matters)."""
import numpy as np
import pytest

from helpers import from_dev
from revenant_amd import synth
from revenant_amd.dsp import FORMAT_CS16, FORMAT_CU8

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

RATE = 122_880_000


@pytest.mark.parametrize("start", [0, (1 << 27) - 4096, 3 * 3600 * RATE])
def test_render_matches_fp64(dev, start):
    scene = synth.wideband_scene(RATE, 40, 11)
    n = 4096
    out = torch.empty(n, dtype=torch.complex64, device="cuda")
    synth.render_device(dev, scene, RATE, start, n, -50.0, 99, -1, out)
    got = from_dev(out).astype(np.complex128)
    want = synth.render_numpy(scene, RATE, start, n, -50.0, 99)
    # float rounding of a sum of magnitude <= ~1.5: a few 1e-7
    assert np.max(np.abs(got - want)) < 1e-6


@pytest.mark.parametrize("fmt", [FORMAT_CS16, FORMAT_CU8])
def test_quantised_codes(dev, fmt):
    scene = synth.wideband_scene(RATE, 12, 5, level_range=(-20.0, -8.0))
    n = 8192
    bps = 4 if fmt == FORMAT_CS16 else 2
    out = torch.empty(n * bps, dtype=torch.uint8, device="cuda")
    synth.render_device(dev, scene, RATE, 1000, n, -40.0, 3, fmt, out)
    raw = from_dev(out)
    x = synth.render_numpy(scene, RATE, 1000, n, -40.0, 3).astype(np.complex64)
    if fmt == FORMAT_CS16:
        codes = raw.view(np.int16).reshape(-1, 2).astype(np.float64)
        want = np.stack([x.real, x.imag], 1).astype(np.float64) * 32768.0
    else:
        codes = raw.reshape(-1, 2).astype(np.float64)
        want = np.stack([x.real, x.imag], 1).astype(np.float64) * 127.5 + 127.5
    assert np.max(np.abs(codes - np.clip(np.rint(want), codes.min(), codes.max()))) <= 1.0
