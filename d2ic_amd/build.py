"""Build libd2ic_amd.so (the HIP kernels + C ABI) in-tree for gfx950.

Usage: ``python -m d2ic_amd.build`` or ``d2ic_amd.build.build()``.
Every source is compiled with hipcc for gfx950 only; packed-FP32 instruction
selection is disabled for all device code.  ``assets.hip`` is additionally
compiled with ``-ffp-contract=off`` so its float arithmetic is rounded exactly
as written (its integer outputs must match the reference bit for bit).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "libd2ic_amd.so"
ARCH = os.environ.get("D2IC_OFFLOAD_ARCH", "gfx950")

SOURCES = {
    "assets.hip": ["-ffp-contract=off"],
    "solver.hip": [],
    "post.hip": [],
    "adjoint.hip": [],
    "local_gn.hip": [],
}

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{PKG.parent / 'include'}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the d2ic_amd HIP library cannot be built")


def _compile(src: str, extra) -> Path:
    obj = OUT_DIR / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
