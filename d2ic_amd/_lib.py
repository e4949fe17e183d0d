"""ctypes binding of libd2ic_amd.so (the C ABI declared in include/d2ic_amd.h).

The library is loaded from the package directory.  There is no fallback: if
the HIP library is missing or the GPU is unavailable, every device entry point
raises, so a silent CPU path can never stand in for the kernels.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libd2ic_amd.so"

D2IC_OK = 0
D2IC_ERR_INVALID = 1
D2IC_ERR_NOMEM = 2
D2IC_ERR_DEVICE = 3
D2IC_ERR_STATE = 4
D2IC_ERR_TIMEOUT = 5
INTERP = {"cubic": 0, "linear": 1}

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (all return int status codes unless listed in _RESTYPE)
_SIGS = {
    "d2ic_version": [],
    "d2ic_pixel_map": [_vp, _i64, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, ctypes.POINTER(_i64), _vp],
    "d2ic_pixel_fill": [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "d2ic_node_pixel_degree": [_vp, _i64, _i64, _vp, ctypes.POINTER(_i32), _vp],
    "d2ic_node_pixel_table": [_vp, _vp, _i64, _i64, _i32, _vp, _vp, _vp],
    "d2ic_cg_plan_create": [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i32, ctypes.POINTER(_vp), _vp],
    "d2ic_cg_plan_create_tiled": [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i32, _i32, _i32,
                                  ctypes.POINTER(_vp), _vp],
    "d2ic_cg_plan_destroy": [_vp],
    "d2ic_cg_set_reference": [_vp, _vp, _i32, _i32, _i32, _vp],
    "d2ic_cg_plan_info": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                          ctypes.POINTER(_i64)],
    "d2ic_cg_solve": [_vp, _vp, _i32, _i32, _vp, _f32, _i32, _f32, _vp, _vp, _vp, _vp],
    "d2ic_cg_solve_batch": [_i32, _pp, _pp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), _pp,
                            ctypes.POINTER(_f32), ctypes.POINTER(_i32), ctypes.POINTER(_f32), _pp, _pp, _pp, _vp],
    "d2ic_cg_solve_batch_noalloc": [_i32, _pp, _pp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), _pp,
                                    ctypes.POINTER(_f32), ctypes.POINTER(_i32), ctypes.POINTER(_f32), _pp, _pp, _pp,
                                    _vp],
    "d2ic_cg_plan_set_grid": [_vp, _i32],
    "d2ic_cg_kernel_ms": [_vp, ctypes.POINTER(_f32)],
    "d2ic_serve_create": [_i32, ctypes.c_double, ctypes.POINTER(_vp)],
    "d2ic_serve_submit": [_vp, _vp, _vp, _i32, _i32, _vp, _vp, _f32, _i32, _f32, _vp, _vp, _vp,
                          ctypes.POINTER(_i32), ctypes.POINTER(_i64)],
    "d2ic_serve_poll": [_vp, _i32, _i64, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)],
    "d2ic_serve_destroy": [_vp],
    "d2ic_adj_sizes": [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)],
    "d2ic_adj_linearize": [_vp, _vp, _i32, _i32, _vp, _i32, _vp, _vp, _vp],
    "d2ic_adj_hvp": [_vp, _vp, _f32, _vp, _i32, _vp, _vp, _vp],
    "d2ic_adj_tangent": [_vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp],
    "d2ic_adj_image_grads": [_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "d2ic_adj_spring": [_vp, _vp, _i32, _vp, _vp, _vp],
    "d2ic_adj_diag_inv": [_vp, _vp, _f32, _vp, _vp],
    "d2ic_lgn_solve": [_vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _i32,
                       _i32, _i32, _vp, _i32, _f32, _f32, _f32, _f32, _vp, _vp],
    "d2ic_lgn_solve_plan": [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _i32, _i32, _i32, _vp, _i32, _f32, _f32, _f32, _f32,
                            _vp, _vp],
    "d2ic_zncc_match": [_vp, _i64, _vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp],
    "d2ic_idw": [_vp, _i64, _vp, _vp, _i64, _vp, _vp],
    "d2ic_strain": [_vp, _vp, _i64, _vp, _vp, _i32, _f32, _f32, _vp, _vp, _vp, _vp],
    "d2ic_discrepancy_map": [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _vp, ctypes.POINTER(_f32), _vp],
    "d2ic_render_speckle": [_vp, _i64, _i32, _i32, _vp, _vp],
}
EXPORTS = ["d2ic_last_error", *_SIGS.keys()]

_lib = None


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first (python -m d2ic_amd.build)")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.d2ic_last_error.restype = ctypes.c_char_p
        lib.d2ic_last_error.argtypes = []
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == D2IC_OK:
        return
    msg = load().d2ic_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == D2IC_ERR_INVALID:
        raise ValueError(text)
    raise RuntimeError(text)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


# ---------------------------------------------------------------------------
# torch plumbing (device memory + streams only)
# ---------------------------------------------------------------------------

def torch():
    import torch as _t

    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("d2ic_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    load()
    return t


def device():
    return torch().device("cuda", torch().cuda.current_device())


def stream() -> int:
    return int(torch().cuda.current_stream().cuda_stream)


def ptr(t) -> int:
    return int(t.data_ptr()) if t is not None else 0


def to_dev(a, dtype, shape=None):
    """Contiguous device tensor of the given torch dtype (copies only if needed)."""
    t = torch()
    if isinstance(a, t.Tensor):
        x = a.to(device=device(), dtype=dtype)
    else:
        x = t.as_tensor(np.asarray(a), device=device()).to(dtype)
    x = x.contiguous()
    if shape is not None:
        x = x.reshape(shape)
    return x


def to_np(a) -> np.ndarray:
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)
