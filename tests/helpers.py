"""Shared test plumbing: seeded inputs and the bit-exact diff.

The diff follows tests/reference/reference_diff.h:98-152 of the reference:
compare bit patterns (so -0.0 != +0.0 and NaN payloads count), and on failure
name the seed, the first divergent index and both patterns.
"""
from __future__ import annotations

import numpy as np

# Seeds in the reference's style (test_pfb.cpp:231-386 uses 0x50464200000000xx).
SEED_BRANCH = 0x5046420000000001
SEED_FFT = 0x5046420000000002
SEED_SPECTRUM = 0x5350454300000001
SEED_CHAIN = 0x4348414E00000001


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def uniform_complex(seed: int, n: int, scale: float = 1.0) -> np.ndarray:
    """float32 pairs uniform in [-scale, scale) (SeededInput, reference_diff.h:185)."""
    return (rng(seed).uniform(-scale, scale, size=2 * n)).astype(np.float32)


def adversarial(seed: int, n: int) -> np.ndarray:
    """Values that stress rounding and flushing (reference_diff.h:216-233):
    denormals, +-1e38, 2^23/2^24 boundaries, signed zeros, mixed with uniform."""
    g = rng(seed)
    base = g.uniform(-1.0, 1.0, size=2 * n).astype(np.float32)
    specials = np.array([0.0, -0.0, 1e-40, -1e-40, 1.17549435e-38, 1e38, -1e38, 8388608.0,
                         8388609.0, 16777216.0, 16777217.0, -16777215.0, 1.0, -1.0, 0.5,
                         3.4e38], dtype=np.float32)
    pick = g.integers(0, 4, size=base.size) == 0
    base[pick] = specials[g.integers(0, specials.size, size=int(pick.sum()))]
    return base


def bits(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a))
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    assert a.dtype == np.float32, a.dtype
    return a.view(np.uint32).ravel()


def assert_bit_exact(gpu, cpu, seed: int | None = None, what: str = "") -> None:
    """Bit-identical, except that any NaN equals any NaN: an invalid operation's
    NaN has a vendor-chosen sign and payload (x86 gives 0xFFC00000, gfx950
    0x7FC00000), and the reference's harness compares NaNs the same way
    (tests/reference/reference_diff.h:71-75). Signed zeros still differ."""
    ga, ca = np.asarray(gpu), np.asarray(cpu)
    if ga.dtype == np.float64 or ca.dtype == np.float64:
        assert ga.dtype == ca.dtype == np.float64, (ga.dtype, ca.dtype)
        g64, c64 = ga.view(np.uint64).ravel(), ca.view(np.uint64).ravel()
        assert g64.size == c64.size, f"{what}: sizes differ {g64.size} vs {c64.size}"
        nan = np.isnan(ga.ravel()) & np.isnan(ca.ravel())
        bad = np.nonzero((g64 != c64) & ~nan)[0]
        if bad.size:
            i = int(bad[0])
            raise AssertionError(f"{what}: {bad.size} of {g64.size} doubles differ; first at "
                                 f"{i}: gpu {ga.ravel()[i]!r} vs twin {ca.ravel()[i]!r}")
        return
    g, c = bits(gpu), bits(cpu)
    assert g.size == c.size, f"{what}: sizes differ {g.size} vs {c.size}"
    nan_g = (g & 0x7F800000) == 0x7F800000
    nan_g &= (g & 0x007FFFFF) != 0
    nan_c = (c & 0x7F800000) == 0x7F800000
    nan_c &= (c & 0x007FFFFF) != 0
    bad = np.nonzero((g != c) & ~(nan_g & nan_c))[0]
    if bad.size:
        i = int(bad[0])
        gf = g[i:i + 1].view(np.float32)[0]
        cf = c[i:i + 1].view(np.float32)[0]
        tag = f" (seed {seed:#x})" if seed is not None else ""
        raise AssertionError(
            f"{what}{tag}: {bad.size} of {g.size} words differ; first at float {i}: "
            f"gpu {gf!r} (0x{g[i]:08x}) vs twin {cf!r} (0x{c[i]:08x})")


def to_dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_dev(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()
