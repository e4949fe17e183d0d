"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* The C oracle, built with gcc ``-fsanitize=address,undefined`` (``make -C
  oracle sanitized``), runs the oracle test files.
* The library's host code, built by hipcc with the sanitizers on the host side
  only (``d2ic_amd.build.build(sanitized=True)``; device code unchanged), runs
  every C-ABI entry point with NULL arrays (``tests/_entry_sweep.py``): each
  must reject them with D2IC_ERR_INVALID before any device call (destroy of
  NULL is a no-op).  The plain library must answer the same.

Each run is a child process that loads the sanitizer runtime first and sees
no device."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE_TESTS = ["tests/test_oracle_assets.py", "tests/test_oracle_objective.py", "tests/test_oracle_sampling.py",
                "tests/test_oracle_zncc_strain_idw.py", "tests/test_oracle_nlcg.py", "tests/test_golden.py",
                "tests/test_oracle_local_gn.py", "tests/test_oracle_discrepancy.py"]
SAN_REPORT = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _env(runtime: str) -> dict:
    env = dict(os.environ)
    pre = env.get("LD_PRELOAD", "")
    env["LD_PRELOAD"] = runtime + (":" + pre if pre else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["HIP_VISIBLE_DEVICES"] = "-1"
    return env


def _gcc_asan() -> str:
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    path = r.stdout.strip()
    if not Path(path).is_absolute():
        pytest.skip("gcc has no shared AddressSanitizer runtime")
    return path


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitized"], check=True)
    env = _env(_gcc_asan())
    env["D2IC_ORACLE_SANITIZED"] = "1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        *ORACLE_TESTS], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert not any(s in out for s in SAN_REPORT), out[-4000:]
    assert r.returncode == 0, out[-4000:]


def _sweep(lib: Path, env: dict) -> dict:
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "_entry_sweep.py"), str(lib)], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert not any(s in out for s in SAN_REPORT), out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return {ln.split()[0]: int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith("d2ic_")}


def _expected(name: str) -> int:
    return 0 if name in ("d2ic_cg_plan_destroy", "d2ic_serve_destroy") else 1


def test_entry_points_reject_null_arrays():
    from d2ic_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    got = _sweep(_lib.LIB_PATH, env)
    assert set(got) == set(_lib._SIGS) - {"d2ic_version"}
    bad = {k: v for k, v in got.items() if v != _expected(k)}
    assert not bad, bad


def test_host_library_under_sanitizers():
    from d2ic_amd import build

    srcs = [*build.CSRC.glob("*"), *(ROOT / "include").glob("*.h"), Path(build.__file__)]
    fresh = build.SAN_LIB.exists() and build.SAN_LIB.stat().st_mtime > max(p.stat().st_mtime for p in srcs)
    try:
        runtime = str(build.sanitizer_runtime())
        lib = build.SAN_LIB if fresh else build.build(verbose=False, sanitized=True)
    except RuntimeError as e:
        pytest.skip(f"sanitized build unavailable: {e}")
    got = _sweep(lib, _env(runtime))
    bad = {k: v for k, v in got.items() if v != _expected(k)}
    assert not bad, bad
