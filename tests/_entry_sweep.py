"""Calls every C-ABI entry point with NULL for each array and handle argument
(sizes positive, scalars valid) and prints ``name rc`` per line.  Run as a
child process by ``tests/test_host_sanitized.py``; argv[1] is the library.
NULL arrays must be rejected before any device call, so no device is used."""
from __future__ import annotations

import ctypes
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from d2ic_amd import _lib as L  # noqa: E402

# int32 arguments of the asset entry points: image 4x4, bounding box [0,4)x[0,4)
_INTS = {
    "d2ic_pixel_map": [4, 4, 0, 4, 0, 4],
    "d2ic_pixel_fill": [4, 4, 0, 4, 0, 4],
}


def main(path: str) -> None:
    lib = ctypes.CDLL(path)
    for name, sig in L._SIGS.items():
        if name == "d2ic_version":
            continue
        ints = list(_INTS.get(name, []))
        args = []
        for t in sig:
            if t is L._i32:
                args.append(ctypes.c_int32(ints.pop(0) if ints else 4))
            elif t is L._i64:
                args.append(ctypes.c_int64(16))
            elif t is L._f32:
                args.append(ctypes.c_float(0.5))
            elif t is ctypes.c_double:
                args.append(ctypes.c_double(1.0))
            else:  # every pointer: NULL
                args.append(None)
        f = getattr(lib, name)
        f.argtypes = sig
        print(name, int(f(*args)), flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
