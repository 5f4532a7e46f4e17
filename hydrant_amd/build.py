"""Build recipe for the native libraries.

``build_native()`` compiles every HIP/C++ source under ``hydrant_amd/csrc`` for
gfx950 with hipcc into ``hydrant_amd/libhydrant_hip.so`` (the C-ABI library
declared in ``include/hydrant_hip.h``).  Objects go to ``hydrant_amd/_build``.

Numerics flags: ``-ffp-contract=off`` keeps every float operation a single
binary32 operation (no FMA contraction), matching the reference's CPU
backend; division and sqrt stay correctly rounded.  Packed-FP32 instructions
are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
REPO = ROOT.parent
CSRC = ROOT / "csrc"
BUILD = ROOT / "_build"
LIB = ROOT / "libhydrant_hip.so"

ARCH = "gfx950"

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{REPO / 'include'}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libhydrant_hip.so")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + src.suffix.replace(".", "_") + ".o")
    cmd = [_hipcc(), *COMMON, *DEVICE, "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build_native(verbose: bool = False) -> Path:
    """Compile all sources and link libhydrant_hip.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), *DEVICE, "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = REPO / "oracle"
    res = subprocess.run(["make", "-C", str(odir)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    out = odir / "_build" / "liboracle.so"
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build_native(verbose=True)
    build_oracle(verbose=True)
    sys.exit(0)
