"""Pixel-level assets built on the GPU.

Drop-in for ``build_pixel_assets`` (reference ``d2ic/pixel_assets.py:21-112``).
The pixel->element first match, the Q4 shape functions (10 Newton steps) and
the node->pixel table are computed by HIP kernels (``d2ic_pixel_map``,
``d2ic_pixel_fill``, ``d2ic_node_pixel_degree``, ``d2ic_node_pixel_table``);
the small regularisation graph (O(Ne) work) is built on the host with numpy
exactly as the reference does.  Outputs are device tensors with the
reference's dtypes (float32 / int32).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .mesh_assets import Mesh, PixelAssets


def _bbox_rect(nodes: np.ndarray, H: int, W: int, pad: float):
    """Candidate rectangle of ``_sample_pixels_in_bbox`` (pixel_assets.py:115-132)."""
    x_min = float(nodes[:, 0].min())
    x_max = float(nodes[:, 0].max())
    y_min = float(nodes[:, 1].min())
    y_max = float(nodes[:, 1].max())
    j0 = max(0, int(math.floor(x_min - pad)))
    j1 = min(W, int(math.ceil(x_max + pad)))
    i0 = max(0, int(math.floor(y_min - pad)))
    i1 = min(H, int(math.ceil(y_max + pad)))
    return i0, max(i0, i1), j0, max(j0, j1)


def build_node_neighbor_dense(elements, nodes_coord, n_nodes):
    """Regularisation graph (pixel_assets.py:330-409): every ordered pair of
    distinct nodes of an element, deduplicated and sorted by (src, dst); the
    slot of an edge is its rank within its source; weight 1/(|x_j - x_i| + 1e-6)
    in the node dtype.  Returns numpy (index (Nn,maxdeg) padded -1, degree,
    weight padded 0)."""
    el = np.asarray(elements, dtype=np.int64)
    X = np.asarray(nodes_coord)
    n_nodes = int(n_nodes)
    empty = (np.full((n_nodes, 0), -1, np.int32), np.zeros((n_nodes,), np.int32), np.zeros((n_nodes, 0), X.dtype))
    if el.size == 0 or n_nodes == 0:
        return empty
    nen = el.shape[1]
    a, b = np.meshgrid(np.arange(nen), np.arange(nen), indexing="ij")
    off = a != b
    src = el[:, a[off]].reshape(-1)
    dst = el[:, b[off]].reshape(-1)
    if src.size == 0:
        return empty
    # lexicographic (src, dst) order = ascending src * n_nodes + dst (one
    # int64 key per edge sorts far faster than row-wise unique)
    key = np.unique(src * n_nodes + dst)
    s, d = np.divmod(key, n_nodes)
    deg = np.bincount(s, minlength=n_nodes).astype(np.int32)
    maxdeg = int(deg.max(initial=0))
    if maxdeg == 0:
        return empty
    dist = np.linalg.norm(X[d, :2] - X[s, :2], axis=1)
    w = (1.0 / (dist + 1e-6)).astype(X.dtype, copy=False)
    first = np.concatenate([[0], np.cumsum(deg)[:-1]])
    slot = np.arange(s.shape[0]) - first[s]
    idx = np.full((n_nodes, maxdeg), -1, np.int32)
    wt = np.zeros((n_nodes, maxdeg), X.dtype)
    idx[s, slot] = d.astype(np.int32)
    wt[s, slot] = w
    return idx, deg, wt


def build_pixel_assets(mesh: Mesh, ref_image, binning: float, roi_mask=None, bbox_pad: float = 2.0,
                       chunk_size: int = 4096) -> PixelAssets:
    """Build :class:`PixelAssets` on the GPU (see module docstring).

    ``chunk_size`` is accepted for signature compatibility; the device search
    does not chunk."""
    t = _lib.require_gpu()
    nodes_in = _lib.to_np(mesh.nodes_xy)
    nodes = (nodes_in / float(binning))
    elements = _lib.to_np(mesh.elements).astype(np.int32)
    shape = tuple(ref_image.shape[:2]) if hasattr(ref_image, "shape") else np.asarray(ref_image).shape[:2]
    H, W = int(shape[0]), int(shape[1])
    roi_d = None
    if roi_mask is not None:
        roi = np.asarray(_lib.to_np(roi_mask), dtype=bool)
        if roi.shape != (H, W):
            raise ValueError(f"roi_mask shape {roi.shape} does not match ref_image shape {(H, W)}")
        roi_d = t.as_tensor(roi.astype(np.uint8), device=_lib.device())
    nn = int(nodes.shape[0])
    ne = int(elements.shape[0])
    nodes32 = nodes.astype(np.float32)
    dev = _lib.device()
    nodes_d = t.as_tensor(nodes32, device=dev).contiguous()
    elems_d = t.as_tensor(elements.reshape(-1, 4) if ne else np.zeros((0, 4), np.int32), device=dev).contiguous()
    if nn > 0:
        i0, i1, j0, j1 = _bbox_rect(nodes, H, W, float(bbox_pad))
    else:
        i0 = i1 = j0 = j1 = 0
    ncand = (i1 - i0) * (j1 - j0)
    cand = t.empty(max(ncand, 1), dtype=t.int32, device=dev)
    n_kept = ctypes.c_int64(0)
    s = _lib.stream()
    _lib.call("d2ic_pixel_map", _lib.ptr(nodes_d), nn, _lib.ptr(elems_d), ne, H, W, i0, i1, j0, j1,
              _lib.ptr(roi_d), _lib.ptr(cand), ctypes.byref(n_kept), s)
    npx = int(n_kept.value)
    coords = t.empty((npx, 2), dtype=t.float32, device=dev)
    pnodes = t.empty((npx, 4), dtype=t.int32, device=dev)
    shapeN = t.empty((npx, 4), dtype=t.float32, device=dev)
    pelts = t.empty((npx,), dtype=t.int32, device=dev)
    roi_flat = t.empty((npx,), dtype=t.int32, device=dev)
    if npx > 0:
        _lib.call("d2ic_pixel_fill", _lib.ptr(cand), H, W, i0, i1, j0, j1, _lib.ptr(nodes_d), _lib.ptr(elems_d),
                  npx, _lib.ptr(coords), _lib.ptr(pnodes), _lib.ptr(shapeN), _lib.ptr(pelts), _lib.ptr(roi_flat), s)
    nb_idx, nb_deg, nb_w = build_node_neighbor_dense(elements, nodes32, nn)
    degree = t.zeros((nn,), dtype=t.int32, device=dev)
    maxdeg = ctypes.c_int32(0)
    if nn > 0:
        _lib.call("d2ic_node_pixel_degree", _lib.ptr(pnodes), npx, nn, _lib.ptr(degree), ctypes.byref(maxdeg), s)
    md = int(maxdeg.value)
    npi = t.empty((nn, md), dtype=t.int32, device=dev)
    npw = t.empty((nn, md), dtype=t.float32, device=dev)
    if nn * md > 0:
        _lib.call("d2ic_node_pixel_table", _lib.ptr(pnodes), _lib.ptr(shapeN), npx, nn, md, _lib.ptr(npi),
                  _lib.ptr(npw), s)
    return PixelAssets(
        pixel_coords_ref=coords,
        pixel_nodes=pnodes,
        pixel_shapeN=shapeN,
        node_neighbor_index=t.as_tensor(nb_idx, device=dev),
        node_neighbor_degree=t.as_tensor(nb_deg, device=dev),
        node_neighbor_weight=t.as_tensor(nb_w.astype(np.float32), device=dev),
        node_reg_weight=t.ones((nn,), dtype=t.float32, device=dev),
        node_pixel_index=npi,
        node_N_weight=npw,
        node_pixel_degree=degree,
        roi_mask_flat=roi_flat,
        image_shape=(H, W),
        pixel_elements=pelts,
    )
