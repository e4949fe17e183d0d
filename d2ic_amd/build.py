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
# host-side AddressSanitizer + UBSan build of the same sources (device code is
# unchanged); for the CPU tests of the entry points' host code only
SAN_DIR = PKG / "_build_san"
SAN_LIB = SAN_DIR / "libd2ic_amd_san.so"
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
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


def _compile(src: str, extra, out_dir: Path = OUT_DIR) -> Path:
    obj = out_dir / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, sanitized: bool = False) -> Path:
    out_dir, lib = (SAN_DIR, SAN_LIB) if sanitized else (OUT_DIR, LIB)
    extra = SAN_FLAGS if sanitized else []
    out_dir.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], [*kv[1], *extra], out_dir), SOURCES.items()))
    tmp = lib.with_suffix(".so.tmp")
    link_extra = ["-shared-libsan", *SAN_FLAGS] if sanitized else []
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *link_extra, "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


def sanitizer_runtime() -> Path:
    """The clang AddressSanitizer runtime the sanitized library links against
    (a process must load it first, e.g. through LD_PRELOAD)."""
    root = Path(hipcc()).resolve().parent.parent
    for base in (root / "llvm" / "lib" / "clang", root / "lib" / "llvm" / "lib" / "clang",
                 Path("/opt/rocm/llvm/lib/clang")):
        hits = sorted(base.glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
        if hits:
            return hits[-1]
    raise RuntimeError("clang AddressSanitizer runtime not found")


if __name__ == "__main__":
    build()
    sys.exit(0)
