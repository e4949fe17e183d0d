"""Mesh / asset containers and host-side mesh tables.

Mirrors ``d2ic/mesh_assets.py:12-164``.  ``Mesh`` holds host numpy arrays
(nodes as float32, like the reference's default-precision JAX arrays, and
int32 connectivity).  ``PixelAssets`` holds device tensors built by
:func:`d2ic_amd.pixel_assets.build_pixel_assets`.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .types import Array


@dataclass(frozen=True)
class Mesh:
    """2-D mesh in image coordinates: x = column, y = row (``mesh_assets.py:12-23``)."""
    nodes_xy: Array  # (Nn, 2)
    elements: Array  # (Ne, 4)


@dataclass(frozen=True)
class MeshAssets:
    """Precomputations derived from a Mesh (``mesh_assets.py:26-50``)."""
    mesh: Mesh
    element_centers_xy: Array  # (Ne, 2)
    node_neighbor_index: Optional[Array] = None
    node_neighbor_degree: Optional[Array] = None
    pixel_data: "PixelAssets | None" = None
    roi_mask: Optional[Array] = None


@dataclass(frozen=True)
class PixelAssets:
    """Pixel-level caches (``mesh_assets.py:55-75``), device-resident.

    ``pixel_elements`` is an addition: the element index of each kept pixel
    (the value the reference computes and then discards,
    ``pixel_assets.py:76-79``).  It is None for assets built elsewhere.
    """
    pixel_coords_ref: Array
    pixel_nodes: Array
    pixel_shapeN: Array
    node_neighbor_index: Array
    node_neighbor_degree: Array
    node_neighbor_weight: Array
    node_reg_weight: Array
    node_pixel_index: Array
    node_N_weight: Array
    node_pixel_degree: Array
    roi_mask_flat: Array
    image_shape: tuple
    pixel_elements: Optional[Array] = None


def _np(a) -> np.ndarray:
    if hasattr(a, "detach"):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def compute_element_centers(mesh: Mesh) -> np.ndarray:
    """Mean of the element's node coordinates, float32 (``mesh_assets.py:77-92``)."""
    nodes = _np(mesh.nodes_xy).astype(np.float32, copy=False)
    el = _np(mesh.elements).astype(np.int64)
    q = nodes[el]  # (Ne, 4, 2)
    s = ((q[:, 0] + q[:, 1]) + q[:, 2]) + q[:, 3]
    return (s / np.float32(q.shape[1])).astype(np.float32)


def build_node_neighbor_tables(mesh: Mesh, max_degree: int = 16):
    """Strain neighbour tables (``mesh_assets.py:95-132``): for each node the
    sorted distinct nodes sharing an element, truncated to ``max_degree`` and
    padded with 0."""
    nodes = _np(mesh.nodes_xy)
    n_nodes = nodes.shape[0]
    el = _np(mesh.elements).astype(np.int64)
    nen = el.shape[1] if el.ndim == 2 else 0
    idx = np.zeros((n_nodes, max_degree), dtype=np.int32)
    deg = np.zeros((n_nodes,), dtype=np.int32)
    if el.size == 0:
        return idx, deg
    src = np.repeat(el, nen, axis=1).reshape(-1)
    dst = np.tile(el, (1, nen)).reshape(-1)
    keep = src != dst
    pairs = np.unique(np.stack([src[keep], dst[keep]], axis=1), axis=0)  # sorted by (src, dst)
    if pairs.size == 0:
        return idx, deg
    s = pairs[:, 0]
    counts = np.bincount(s, minlength=n_nodes)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    slot = np.arange(pairs.shape[0]) - starts[s]
    ok = slot < max_degree
    idx[s[ok], slot[ok]] = pairs[ok, 1].astype(np.int32)
    deg[:] = np.minimum(counts, max_degree).astype(np.int32)
    return idx, deg


def make_mesh_assets(mesh: Mesh, with_neighbors: bool = True, max_degree: int = 16) -> MeshAssets:
    """``mesh_assets.py:135-164``."""
    centers = compute_element_centers(mesh)
    node_idx = node_deg = None
    if with_neighbors:
        node_idx, node_deg = build_node_neighbor_tables(mesh, max_degree=max_degree)
    return MeshAssets(mesh=mesh, element_centers_xy=centers, node_neighbor_index=node_idx,
                      node_neighbor_degree=node_deg, pixel_data=None)
