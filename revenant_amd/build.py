"""Build recipes: the HIP library (gfx950) and the CPU oracle.

``build_library()`` compiles every source under ``revenant_amd/csrc`` with
hipcc for gfx950 and links ``revenant_amd/librvx_hip.so``. The numeric flags
are part of the contract the kernels keep with the reference twins:

* ``-ffp-contract=off``: no multiply-add fusion, so each product and sum
  rounds on its own exactly as in the twins (core/dsp/reference_fp.h).
* ``-fgpu-flush-denormals-to-zero``: FP32 denormals flushed on input and
  output, the twins' FTZ|DAZ (core/dsp/denormal_mode.h).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.

``build_oracle()`` runs ``make`` in ``oracle/`` and ``build_negative_control()``
compiles the suite's deliberately wrong kernel (both test infrastructure).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "revenant_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librvx_hip.so"
ORACLE = ROOT / "oracle"
ARCH = os.environ.get("RVX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the library)")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [hipcc(), *HIP_FLAGS, *lang, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def sources() -> list[Path]:
    return sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    res = subprocess.run(["make", "-C", str(ORACLE), "-s", "liboracle.so"], capture_output=True,
                         text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return ORACLE / "liboracle.so"


NEGATIVE = ROOT / "tests" / "negative_control"
NEGATIVE_LIB = NEGATIVE / "libcmul_wrong.so"


def build_negative_control() -> Path:
    """The conformance suite's deliberately wrong kernel, in a library of its
    own under tests/ (the product library never contains it)."""
    src = NEGATIVE / "cmul_wrong.hip"
    tmp = NEGATIVE_LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-x", "hip", "-shared", str(src), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, NEGATIVE_LIB)
    return NEGATIVE_LIB


def build_all() -> None:
    build_oracle()
    build_library()
    build_negative_control()


if __name__ == "__main__":
    build_all()
    print(LIB)
