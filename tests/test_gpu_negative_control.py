"""A deliberately wrong kernel fails the diff (tests/reference/test_divergence.cpp:84-125).

tests/negative_control/cmul_wrong.hip is the device complex multiply with the
sign of the real part's second term flipped. Its outputs are finite and
plausible; the bit diff against the oracle must reject them, on real parts
only, while the product library's own cmul passes on the same inputs.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from helpers import diff, from_dev, to_dev, uniform_complex

LIB = Path(__file__).resolve().parent / "negative_control" / "libcmul_wrong.so"


def _wrong_lib():
    if not LIB.exists():
        raise RuntimeError(f"{LIB} is not built (revenant_amd.build.build_negative_control)")
    lib = ctypes.CDLL(str(LIB))
    lib.cmul_wrong.restype = ctypes.c_int32
    lib.cmul_wrong.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint64, ctypes.c_void_p]
    return lib


@pytest.mark.gpu
def test_wrong_kernel_fails_the_diff(dev):
    from oracle import oracle

    seed = 0xBADC0DE000000001
    n = 8192
    a = uniform_complex(seed, n)
    b = uniform_complex(seed + 1, n)
    da, db = to_dev(a), to_dev(b)
    out = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    print(f"negative control: cmul_wrong n={n} seed={seed:#x}", flush=True)
    assert _wrong_lib().cmul_wrong(da.data_ptr(), db.data_ptr(), out.data_ptr(), n, stream) == 0
    wrong = from_dev(out)
    right = oracle.cmul(a, b)
    res = diff(wrong, right, seed=seed, what="cmul_wrong")
    assert not res.identical and res.max_ulp_error > 0
    assert "seed" in res.report
    assert res.first_divergence % 2 == 0
    assert diff(wrong[1::2], right[1::2]).identical, "only real parts may diverge"
    assert np.all(np.isfinite(wrong))

    # the product's kernel on the same inputs passes the same diff
    good = torch.zeros(2 * n, dtype=torch.float32, device="cuda")
    dev.cmul(da, db, good)
    assert diff(from_dev(good), right, seed=seed, what="cmul").identical
