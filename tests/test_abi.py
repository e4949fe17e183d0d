"""The C-ABI library loads without a GPU and exports every entry point that
include/d2ic_amd.h declares; the Python layer's ctypes signature table names
exactly those entry points.  No compute call is made (no GPU here)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "d2ic_amd.h"
LIB = ROOT / "d2ic_amd" / "libd2ic_amd.so"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    return re.findall(r"^[A-Za-z_][\w \*]*?\b(d2ic_\w+)\s*\(", text, flags=re.M)


def test_header_declarations_parse():
    names = _declared()
    assert len(names) == len(set(names)) >= 25
    for must in ("d2ic_cg_solve", "d2ic_cg_plan_create", "d2ic_pixel_map", "d2ic_zncc_match", "d2ic_discrepancy_map",
                 "d2ic_lgn_solve", "d2ic_adj_hvp", "d2ic_serve_submit"):
        assert must in names


@pytest.mark.skipif(not LIB.exists(), reason="library not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert missing == []
    lib.d2ic_version.restype = ctypes.c_int
    assert lib.d2ic_version() >= 100
    lib.d2ic_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.d2ic_last_error(), bytes)


def test_python_signature_table_matches_header():
    from d2ic_amd import _lib

    assert set(_lib.EXPORTS) == set(_declared())  # d2ic_last_error + the ctypes signature table
