"""Structured mesher and island removal (reference ``d2ic/mask2mesh.py:26-142,
391-432``): the reference's own smoke test (``tests/test_mask2mesh_smoke.py``),
the node/element counts of the benchmark workloads computed from the
reference's rules, and closed forms for the rules a vectorised restatement can
get wrong (half-to-even rounding of cell centres, 4- not 8-connectivity,
block-OR binning with padding, node numbering)."""
from __future__ import annotations

import numpy as np
import pytest

from d2ic_amd import speckle
from d2ic_amd.mask2mesh import _remove_small_components, downsample_mask, mask_to_mesh


def test_reference_smoke_filters_small_islands():
    """tests/test_mask2mesh_smoke.py:8-25 of the reference."""
    mask = np.zeros((64, 64), dtype=bool)
    mask[8:48, 10:50] = True
    mask[2, 2] = True  # tiny island
    mesh = mask_to_mesh(mask, element_size_px=8, binning=1, remove_islands=True, min_island_area_px=10)
    assert mesh.nodes_xy.shape[1] == 2 and mesh.nodes_xy.shape[0] > 0
    assert mesh.elements.shape[1] == 4 and mesh.elements.shape[0] > 0
    centers = (np.asarray(mesh.nodes_xy)[np.asarray(mesh.elements)].mean(axis=1)).astype(int)
    assert not np.any((centers[:, 0] == 2) & (centers[:, 1] == 2))
    # the island is no cell centre (centres sit at 4 + 8k), so the mesh is the same without removal
    mesh2 = mask_to_mesh(mask, element_size_px=8, remove_islands=False)
    assert mesh2.elements.shape[0] == mesh.elements.shape[0]
    # a 9-pixel island over the centre (4, 4) of cell (0, 0): kept as an element
    # unless island removal (area < 10) drops it first
    mask[3:6, 3:6] = True
    n_keep = mask_to_mesh(mask, 8, remove_islands=False).elements.shape[0]
    n_drop = mask_to_mesh(mask, 8, remove_islands=True, min_island_area_px=10).elements.shape[0]
    assert n_keep == n_drop + 1 == mesh.elements.shape[0] + 1


@pytest.mark.parametrize("case, nn, ne", [("cfg0", 1089, 625), ("cfg1", 16641, 11430), ("cfg3", 4225, 2704)])
def test_benchmark_mesh_counts(case, nn, ne):
    if case == "cfg0":
        m = speckle.rect_mask(512, 512, 400, 400)
    elif case == "cfg1":
        m = speckle.plate_with_hole_mask(2048, 2048, 1800, 300.0)
    else:
        # 2704 = 52^2 cells: the centred square whose cell centres 8 + 16k
        # span k = 6..57 is 832 px wide (an 800 px square keeps 50^2)
        m = speckle.rect_mask(1024, 1024, 832, 832)
        assert mask_to_mesh(speckle.rect_mask(1024, 1024, 800, 800), 16).elements.shape[0] == 2500
    mesh = mask_to_mesh(m, 16)
    assert mesh.nodes_xy.shape[0] == nn and mesh.elements.shape[0] == ne


def test_cell_centre_rounds_half_to_even():
    # step 15 on width 64: axis 0, 15, 30, 45, 60, 63; cell centres 7.5, 22.5,
    # 37.5, 52.5, 61.5 round (half to even) to 8, 22, 38, 52, 62.  Only column
    # 22 is set: the second cell column is kept (floor(c + .5) would test 23).
    mask = np.zeros((64, 64), bool)
    mask[:, 22] = True
    mesh = mask_to_mesh(mask, 15, remove_islands=False)
    nodes = np.asarray(mesh.nodes_xy)
    np.testing.assert_array_equal(np.unique(nodes[:, 0]), [0, 15, 30, 45, 60, 63])
    x0 = nodes[np.asarray(mesh.elements)[:, 0], 0]
    np.testing.assert_array_equal(np.unique(x0), [15])
    assert mesh.elements.shape[0] == 5  # every cell row of that column
    mask2 = np.zeros((64, 64), bool)
    mask2[:, 23] = True
    with pytest.raises(ValueError):
        mask_to_mesh(mask2, 15, remove_islands=False)


def test_islands_are_four_connected():
    # two 40-pixel blocks touching only at a corner are two components: both
    # fall below min area 64 (one 8-connected component of 80 would stay)
    m = np.zeros((40, 40), bool)
    m[0:8, 0:5] = True
    m[8:16, 5:10] = True
    m[20:40, 20:40] = True  # 400 px, kept
    out = _remove_small_components(m, 64)
    assert not out[0:16, 0:10].any()
    assert out[20:40, 20:40].all() and out.sum() == 400
    np.testing.assert_array_equal(_remove_small_components(m, 40), m)
    np.testing.assert_array_equal(_remove_small_components(np.zeros((5, 5), bool), 64), np.zeros((5, 5), bool))


def test_binning_block_or_with_padding_and_node_numbering():
    m = np.zeros((10, 10), bool)
    m[9, 9] = True
    d = downsample_mask(m, 3)
    assert d.shape == (4, 4) and d[3, 3] and d.sum() == 1
    mask = speckle.rect_mask(96, 80, 64, 48)
    mesh = mask_to_mesh(mask, 16)
    nodes = np.asarray(mesh.nodes_xy)
    el = np.asarray(mesh.elements)
    xs = np.unique(nodes[:, 0])
    nx = xs.size
    # n = iy * nx + ix with x fastest; element [n0, n0 + 1, n0 + nx + 1, n0 + nx]
    iy, ix = np.divmod(np.arange(nodes.shape[0]), nx)
    np.testing.assert_array_equal(nodes[:, 0], xs[ix])
    np.testing.assert_array_equal(el[:, 1] - el[:, 0], 1)
    np.testing.assert_array_equal(el[:, 2] - el[:, 0], nx + 1)
    np.testing.assert_array_equal(el[:, 3] - el[:, 0], nx)
    assert np.all(np.diff(el[:, 0]) > 0)  # iy-major, ix-minor order
    with pytest.raises(ValueError):
        mask_to_mesh(np.zeros((32, 32), bool), 8)
    with pytest.raises(ValueError):
        mask_to_mesh(mask, 0)
    with pytest.raises(NotImplementedError):
        mask_to_mesh(mask, 16, enforce_quads=False)
