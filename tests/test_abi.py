"""The C ABI library loads without a GPU and exports every symbol include/rvx.h
declares; host-side design entry points work on the CPU; launches without a
device fail with a message rather than crash."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from revenant_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared() -> set[str]:
    text = (ROOT / "include" / "rvx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(rvx_[a-z0-9_]+)\s*\(", text))


def test_every_declared_symbol_is_exported():
    lib = _lib.lib()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.exported_symbols()) <= names


def test_abi_version_and_struct_sizes():
    assert _lib.lib().rvx_abi_version() == 1
    assert C.sizeof(_lib.VrxFineParams) == 15 * 4
    assert C.sizeof(_lib.VrxDemodParams) == 4 * 4
    assert C.sizeof(_lib.PfbFftParams) == 7 * 4
    assert C.sizeof(_lib.SpectrumParams) == 6 * 4
    # the noise stages' parameter blocks, word for word the reference's push constants
    assert C.sizeof(_lib.BlankerParams) == 7 * 4      # noise_reference.h:165
    assert C.sizeof(_lib.LineParams) == 32 * 4        # :381
    assert C.sizeof(_lib.SpectralParams) == 10 * 4    # :514
    assert C.sizeof(_lib.NoiseLineItem) == 144 and C.sizeof(_lib.NoiseSpectralItem) == 64
    assert C.sizeof(_lib.ExcisionParams) == 8 * 4 and C.sizeof(_lib.CombParams) == 8 * 4
    assert C.sizeof(_lib.VrxDisplayPlan) == 88


def test_design_refuses_bad_grids():
    import pytest

    from revenant_amd.dsp import GridParams, design_prototype

    for g in (GridParams(63, 16, 63), GridParams(64, 0, 32), GridParams(64, 16, 48)):
        with pytest.raises(_lib.RvxError):
            design_prototype(g)


def test_nco_delta_exact():
    from revenant_amd.dsp import nco_delta, turn_fixed64

    assert turn_fixed64(1, 2) == 1 << 63
    assert turn_fixed64(1, 4) == 1 << 62
    assert turn_fixed64(1, 3) == (1 << 64) // 3 + (1 if ((1 << 64) % 3) * 2 >= 3 else 0)
    # round(2^64 * f / rate) by exact integer arithmetic, negative wraps
    for num, den, rate in [(12500, 1, 48000), (-200000, 1, 600000), (1, 3, 44100)]:
        got = nco_delta(num, den, rate)
        from fractions import Fraction
        f = Fraction(num, den) / rate
        f -= int(np.floor(float(f)))
        exact = f * (1 << 64)
        assert abs(Fraction(got) - exact) <= Fraction(1, 2), (num, den, rate)


def test_launch_without_device_reports_an_error():
    lib = _lib.lib()
    rc = lib.rvx_stream_sync(None)
    assert rc == -1
    assert b"null context" in lib.rvx_last_error()



def test_prescale_taps_exact_and_refusals():
    import numpy as np

    from revenant_amd.dsp import prescale_taps

    t = np.array([0.0, 1.0, -3.5, 1e-30, 2.0 ** -111, 65504.0], dtype=np.float32)
    out = prescale_taps(t)
    assert np.array_equal(out, t * np.float32(2.0 ** -15))
    for bad in (np.float32(2.0 ** -112), np.float32(np.inf), np.float32(np.nan)):
        with pytest.raises(_lib.RvxError):
            prescale_taps(np.array([1.0, bad], dtype=np.float32))
