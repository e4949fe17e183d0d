"""Reference-domain discrepancy map on the GPU (reference
``d2ic/discrepancy.py:86-127``): r(x) = I_def(x + u(x)) - I_ref(x) with the
bilinear sampler that returns NaN outside [0, H-1) x [0, W-1), scattered onto
the image through ``roi_mask_flat``; NaN outside the ROI.  Returns
(map (H, W) float32 device tensor, rms float)."""
from __future__ import annotations

import ctypes

from . import _lib
from .mesh_assets import PixelAssets


def compute_discrepancy_map_ref(*, ref_image, def_image, u_nodal, pixel_assets: PixelAssets):
    t = _lib.require_gpu()
    ref = _lib.to_dev(ref_image, t.float32)
    de = _lib.to_dev(def_image, t.float32)
    if ref.dim() != 2 or de.dim() != 2:
        raise ValueError("ref_image and def_image must be 2D arrays.")
    H, W = pixel_assets.image_shape
    u = _lib.to_dev(u_nodal, t.float32).reshape(-1, 2)
    xy = _lib.to_dev(pixel_assets.pixel_coords_ref, t.float32)
    pn = _lib.to_dev(pixel_assets.pixel_nodes, t.int32)
    N = _lib.to_dev(pixel_assets.pixel_shapeN, t.float32)
    rf = _lib.to_dev(pixel_assets.roi_mask_flat, t.int32)
    out = t.empty((int(H), int(W)), dtype=t.float32, device=_lib.device())
    rms = ctypes.c_float(0.0)
    _lib.call("d2ic_discrepancy_map", _lib.ptr(ref), _lib.ptr(de), int(H), int(W), _lib.ptr(u), _lib.ptr(xy),
              _lib.ptr(pn), _lib.ptr(N), _lib.ptr(rf), int(xy.shape[0]), _lib.ptr(out), ctypes.byref(rms),
              _lib.stream())
    return out, float(rms.value)
