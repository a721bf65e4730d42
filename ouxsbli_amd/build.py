"""Build recipe of the gfx950 library ``ouxsbli_amd/libouxsbli.so``.

Each translation unit is compiled with hipcc for gfx950 into an object, then
linked into one shared library that exports the C ABI of include/ouxsbli.h.
The FP64 units are compiled with -ffp-contract=off (every multiply and add
rounded separately, in the evaluation order of the reference expressions);
the FP32-bulk units contract a*b+c within one source expression
(-ffp-contract=on), a rule that does not depend on the surrounding code, so
the single-case and the batched launches of a case round identically, take
FP32 division and square root in their fast hardware forms (about 1 ulp
instead of correctly rounded) and flush FP32 denormals to zero (FP64
arithmetic keeps its denormals).  Packed FP32 instructions are disabled for
all device code.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
LIB = PKG / "libouxsbli.so"
BUILD = PKG / "_build"
ARCH = os.environ.get("OX_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    f"-I{ROOT / 'include'}",
]
# FP32-bulk units: contraction within an expression, fast FP32 division and
# square root, FP32 denormals flushed to zero (which lets the fast division be
# a reciprocal and a multiply, without the range scaling denormals need)
FP32_BULK = ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]
# 3-D FP64 stage kernels: the scheduler that groups memory operations into
# clauses (1% on the FP64 step at 256^3, 3% on its z sweep; scheduling leaves
# the arithmetic, and so the results, unchanged)
FP64_STAGE = ["-ffp-contract=off", "-mllvm", "-amdgpu-sched-strategy=max-memory-clause"]
UNITS = {
    "ox_capi.hip": ["-ffp-contract=off"],
    "ox_batch.hip": ["-ffp-contract=off"],
    "ox_fp64.hip": FP64_STAGE,
    "ox_fp64_batch.hip": FP64_STAGE,
    "ox_fp32.hip": FP32_BULK,
    "ox_fp32_batch.hip": FP32_BULK,
    "ox2d_capi.hip": ["-ffp-contract=off"],
    "ox_ic.hip": ["-ffp-contract=off"],
    "ox2d_fp64.hip": ["-ffp-contract=off"],
    "ox2d_fp32.hip": FP32_BULK,
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and pathlib.Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def _compile(unit: str, flags) -> pathlib.Path:
    obj = BUILD / (unit + ".o")
    cmd = [hipcc(), *COMMON, *flags, "-c", str(CSRC / unit), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
