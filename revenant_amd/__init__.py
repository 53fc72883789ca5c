"""MI355X-native receive chain (gfx950 HIP kernels behind a C ABI).

The device operators live in :mod:`revenant_amd.ops` (they import torch); the
host-side design vocabulary in :mod:`revenant_amd.dsp`; the C ABI binding in
:mod:`revenant_amd._lib`; build recipes in :mod:`revenant_amd.build`.
"""

__all__ = ["dsp", "ops", "build"]
