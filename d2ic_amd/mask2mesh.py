"""Structured Q4 mesher from an ROI mask (host preprocessing).

Behaviour-identical restatement of the reference's structured mesher
(``d2ic/mask2mesh.py:26-182, 391-432``): same axis coordinates, node numbering
``n0 = iy*nx + ix``, element keep rule (cell centre rounded half-to-even
inside the downsampled mask) and island removal (4-connected components
smaller than ``min_island_area_px`` dropped).  Island removal uses
scipy.ndimage.label instead of the reference's pure-Python flood fill; the
kept pixel set is the same.  The gmsh-based mesher is out of scope.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from .mesh_assets import Mesh, MeshAssets, compute_element_centers


def downsample_mask(mask: np.ndarray, binning: int) -> np.ndarray:
    """Block-OR downsampling with zero padding (``mask2mesh.py:26-57``)."""
    mask = np.asarray(mask, dtype=bool)
    if binning <= 1:
        return mask
    h, w = mask.shape
    pad_h = (-h) % binning
    pad_w = (-w) % binning
    if pad_h or pad_w:
        mask = np.pad(mask, ((0, pad_h), (0, pad_w)), mode="constant", constant_values=False)
    new_h = mask.shape[0] // binning
    new_w = mask.shape[1] // binning
    return np.any(mask.reshape(new_h, binning, new_w, binning), axis=(1, 3))


def _build_axis_coords(size: int, step: int) -> np.ndarray:
    """``mask2mesh.py:391-397``."""
    count = max(2, int(np.floor((size - 1) / step)) + 1)
    coords = np.arange(count, dtype=np.float64) * step
    if coords[-1] < size - 1:
        coords = np.append(coords, size - 1)
    return coords


def _remove_small_components(mask: np.ndarray, min_area: int) -> np.ndarray:
    """Drop 4-connected components with area < max(1, min_area) (``mask2mesh.py:400-432``)."""
    from scipy import ndimage

    mask = np.asarray(mask, dtype=bool)
    labels, n = ndimage.label(mask, structure=np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]]))
    if n == 0:
        return mask
    area = np.bincount(labels.ravel(), minlength=n + 1)
    keep = area >= max(1, min_area)
    keep[0] = False
    return keep[labels]


def mask_to_mesh(mask: np.ndarray, element_size_px: float, binning: int = 1,
                 origin_xy: Tuple[float, float] = (0.0, 0.0), enforce_quads: bool = True,
                 remove_islands: bool = True, min_island_area_px: int = 64) -> Mesh:
    """Regular grid over the mask; keep cells whose centre is in the ROI (``mask2mesh.py:60-142``)."""
    if not enforce_quads:
        raise NotImplementedError("Only quadrilateral meshes are supported at the moment.")
    if element_size_px <= 0:
        raise ValueError("element_size_px must be positive")
    mask_bool = np.asarray(mask, dtype=bool)
    if remove_islands:
        mask_bool = _remove_small_components(mask_bool, min_island_area_px)
    mask_ds = downsample_mask(mask_bool, binning)
    if not np.any(mask_ds):
        raise ValueError("mask_to_mesh: no ROI pixels remain after preprocessing.")
    step = max(1, int(round(element_size_px / max(1, binning))))
    h, w = mask_ds.shape
    grid_x = _build_axis_coords(w, step)
    grid_y = _build_axis_coords(h, step)
    nx, ny = grid_x.size, grid_y.size
    node_coords = np.stack(np.meshgrid(grid_x, grid_y), axis=-1).reshape(-1, 2).astype(np.float64)
    node_coords[:, 0] = origin_xy[0] + node_coords[:, 0] * binning
    node_coords[:, 1] = origin_xy[1] + node_coords[:, 1] * binning

    cx = 0.5 * (grid_x[:-1] + grid_x[1:])
    cy = 0.5 * (grid_y[:-1] + grid_y[1:])
    mx = np.minimum(np.rint(cx).astype(np.int64), w - 1)  # round half to even, like round()
    my = np.minimum(np.rint(cy).astype(np.int64), h - 1)
    keep = mask_ds[my[:, None], mx[None, :]]  # (ny-1, nx-1), iy-major like the loop
    iy, ix = np.nonzero(keep)
    n0 = iy * nx + ix
    elements = np.stack([n0, n0 + 1, n0 + nx + 1, n0 + nx], axis=1).astype(np.int32)
    if elements.shape[0] == 0:
        raise ValueError("mask_to_mesh: no elements found inside the ROI.")
    nodes = node_coords.astype(np.float32)
    _sanity_check_mesh(nodes, elements)
    return Mesh(nodes_xy=nodes, elements=elements)


def mask_to_mesh_assets(mask: np.ndarray, element_size_px: float, binning: int = 1,
                        origin_xy: Tuple[float, float] = (0.0, 0.0), enforce_quads: bool = True,
                        remove_islands: bool = True, min_island_area_px: int = 64):
    """Mesh plus minimal assets with the (downsampled) ROI mask (``mask2mesh.py:145-182``)."""
    mesh = mask_to_mesh(mask=mask, element_size_px=element_size_px, binning=binning, origin_xy=origin_xy,
                        enforce_quads=enforce_quads, remove_islands=remove_islands,
                        min_island_area_px=min_island_area_px)
    centers = compute_element_centers(mesh)
    mask_bool = np.asarray(mask, dtype=bool)
    if remove_islands:
        mask_bool = _remove_small_components(mask_bool, min_island_area_px)
    mask_ds = downsample_mask(mask_bool, binning)
    return mesh, MeshAssets(mesh=mesh, element_centers_xy=centers, pixel_data=None, roi_mask=mask_ds)


def mask_to_mesh_gmsh(*args, **kwargs):  # pragma: no cover - out of scope
    raise ImportError("mask_to_mesh_gmsh requires gmsh, which is outside this package's scope.")


def mask_to_mesh_assets_gmsh(*args, **kwargs):  # pragma: no cover - out of scope
    raise ImportError("mask_to_mesh_assets_gmsh requires gmsh, which is outside this package's scope.")


def _sanity_check_mesh(nodes, elements) -> None:
    """``mask2mesh.py:435-443``."""
    if np.isnan(np.asarray(nodes)).any():
        raise ValueError("Mesh nodes contain NaNs")
    elem = np.asarray(elements)
    if elem.size == 0 or elem.shape[1] != 4:
        raise ValueError("Elements must be quad connectivity")
    if np.any(elem < 0) or np.any(elem > nodes.shape[0] - 1):
        raise ValueError("Element connectivity out of bounds")
