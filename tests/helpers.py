"""Shared test plumbing: seeded inputs and the bit-exact diff.

The diff follows tests/reference/reference_diff.h:98-152 of the reference:
compare bit patterns (so -0.0 != +0.0), treat a NaN as matching any NaN (its
sign and payload are the vendor's choice), and on failure name the seed, the
first divergent index and both patterns.
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


class DiffResult:
    """The outcome of one comparison (tests/reference/reference_diff.h:45-62):
    whether the buffers agree within the tolerance, how many elements were
    compared, the first index that did not, the worst distance in units in the
    last place and in absolute terms, and a report naming the seed, the index
    and both bit patterns."""

    def __init__(self) -> None:
        self.identical = True
        self.compared = 0
        self.first_divergence = 0
        self.max_ulp_error = 0
        self.max_absolute_error = 0.0
        self.report = ""

    def __bool__(self) -> bool:
        return self.identical


def _flat(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a))
    if a.dtype == np.complex64:
        a = a.view(np.float32)
    elif a.dtype == np.complex128:
        a = a.view(np.float64)
    return a.ravel()


def _ordered(v: np.ndarray) -> np.ndarray:
    """Float bit patterns mapped onto one increasing integer line: +0 is 0, the
    next float up is 1, -0 is -1, so the distance between two keys counts the
    representable values between them and +0 and -0 are one step apart."""
    if v.dtype.itemsize == 4:
        raw = v.view(np.int32).astype(np.int64)
        return np.where(raw < 0, -(raw & 0x7FFFFFFF) - 1, raw)
    raw = v.view(np.int64)
    return np.where(raw < 0, -(raw & np.int64(0x7FFFFFFFFFFFFFFF)) - 1, raw)


def _distance(kg: np.ndarray, kc: np.ndarray) -> np.ndarray:
    """|kg - kc| without int64 overflow: exact when the keys are equal or close
    (below 2^53 apart), rounded beyond that, never 0 for different keys."""
    far = np.abs(kg.astype(np.float64) - kc.astype(np.float64))
    near = np.abs(np.where(np.sign(kg) == np.sign(kc), kg - kc, 0))
    d = np.where(far < 2.0 ** 53, near.astype(np.float64), far)
    return np.where(kg == kc, 0.0, np.maximum(d, 1.0))


def diff(gpu, cpu, seed: int | None = None, tolerance_ulp: int = 0, what: str = "") -> DiffResult:
    """Compare two buffers element by element.

    Floating-point buffers are compared as bit patterns: the distance is the
    number of representable values between the two (so +0 and -0 differ, and
    one rounding decision is 1 ulp); a NaN matches only a NaN, whatever its
    sign and payload, since an invalid operation's NaN is vendor-chosen (x86
    gives 0xFFC00000, gfx950 0x7FC00000). Integer buffers must match exactly.
    ``tolerance_ulp`` of 0 demands bit-identical output; a larger one must be
    asked for by the caller, who says why (reference_diff.h:92-152)."""
    g, c = _flat(gpu), _flat(cpu)
    res = DiffResult()
    if g.dtype != c.dtype:
        res.identical = False
        res.report = f"{what}: dtype mismatch: gpu {g.dtype}, reference {c.dtype}\n"
        return res
    if g.size != c.size:
        res.identical = False
        res.report = (f"{what}: size mismatch: gpu produced {g.size}, "
                      f"reference produced {c.size}\n")
        return res
    res.compared = int(g.size)
    if g.size == 0:
        return res
    if g.dtype.kind == "f":
        nan_g, nan_c = np.isnan(g), np.isnan(c)
        ulp = _distance(_ordered(g), _ordered(c))
        ulp = np.where(nan_g & nan_c, 0.0, ulp)
        ulp = np.where(nan_g ^ nan_c, np.inf, ulp)
        both_finite = ~(nan_g | nan_c)
        with np.errstate(invalid="ignore", over="ignore"):
            absolute = np.where(both_finite, np.abs(g.astype(np.float64) - c.astype(np.float64)),
                                0.0)
        width = 8 if g.dtype.itemsize == 4 else 16
        pattern = lambda x: f"0x{int(np.asarray(x).view(np.uint32 if width == 8 else np.uint64)):0{width}X}"
    else:
        wide = g.astype(np.float64) - c.astype(np.float64)
        ulp = np.where(g == c, 0.0, np.maximum(np.abs(wide), 1.0))
        absolute = np.abs(wide)
        pattern = lambda x: str(x)
    ulp = np.asarray(ulp, dtype=np.float64)
    worst = float(ulp.max())
    res.max_ulp_error = int(worst) if np.isfinite(worst) else worst
    finite_abs = absolute[np.isfinite(absolute)]
    res.max_absolute_error = float(finite_abs.max()) if finite_abs.size else 0.0
    bad = np.nonzero(ulp > tolerance_ulp)[0]
    if bad.size:
        res.identical = False
        res.first_divergence = int(bad[0])
        lines = []
        for i in bad[:8]:
            lines.append(f"  [{int(i)}] gpu={g[i]!r} ({pattern(g[i])})  cpu={c[i]!r} "
                         f"({pattern(c[i])})  ulp={ulp[i]:.0f}")
        tag = f"{seed:#x}" if seed is not None else "none"
        res.report = (f"{what}: GPU output diverges from the oracle\n"
                      f"  seed              {tag}\n"
                      f"  elements compared {res.compared}\n"
                      f"  divergent         {bad.size}\n"
                      f"  first divergence  index {res.first_divergence}\n"
                      f"  worst error       {res.max_ulp_error} ulp, "
                      f"{res.max_absolute_error:.9g} absolute\n"
                      f"  tolerance         {tolerance_ulp} ulp\n" + "\n".join(lines))
    return res


def assert_bit_exact(gpu, cpu, seed: int | None = None, what: str = "") -> None:
    """Raise unless ``diff`` finds the buffers bit-identical (NaN matching any
    NaN, signed zeros distinct)."""
    res = diff(gpu, cpu, seed=seed, what=what)
    if not res.identical:
        raise AssertionError(res.report)


def to_dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_dev(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()
