"""Element centres and strain neighbour tables (reference
``d2ic/mesh_assets.py:77-132``): closed forms on a structured mesh (the 8, 5
or 3 nodes around interior, edge and corner nodes, ascending, zero padded) and
truncation to ``max_degree`` keeping the smallest indices."""
from __future__ import annotations

import numpy as np

from d2ic_amd import speckle
from d2ic_amd.mask2mesh import mask_to_mesh
from d2ic_amd.mesh_assets import Mesh, build_node_neighbor_tables, compute_element_centers, make_mesh_assets


def test_structured_neighbours_and_centres():
    mesh = mask_to_mesh(speckle.rect_mask(64, 80, 64, 80), 16)
    nodes = np.asarray(mesh.nodes_xy)
    xs, ys = np.unique(nodes[:, 0]), np.unique(nodes[:, 1])
    nx, ny = xs.size, ys.size
    assert mesh.elements.shape[0] == (nx - 1) * (ny - 1)  # every cell kept
    idx, deg = build_node_neighbor_tables(mesh, max_degree=16)
    idx, deg = np.asarray(idx), np.asarray(deg)
    for n in range(nodes.shape[0]):
        iy, ix = divmod(n, nx)
        want = sorted((iy + dy) * nx + ix + dx for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                      if (dy or dx) and 0 <= iy + dy < ny and 0 <= ix + dx < nx)
        assert deg[n] == len(want)
        np.testing.assert_array_equal(idx[n, :deg[n]], want)
        assert np.all(idx[n, deg[n]:] == 0)
    assert sorted(set(deg.tolist())) == [3, 5, 8]
    c = np.asarray(compute_element_centers(mesh))
    el = np.asarray(mesh.elements)
    np.testing.assert_array_equal(c[:, 0], 0.5 * (nodes[el[:, 0], 0] + nodes[el[:, 1], 0]))
    np.testing.assert_array_equal(c[:, 1], 0.5 * (nodes[el[:, 0], 1] + nodes[el[:, 3], 1]))
    a = make_mesh_assets(mesh, with_neighbors=True)
    np.testing.assert_array_equal(np.asarray(a.node_neighbor_index), idx)
    assert make_mesh_assets(mesh, with_neighbors=False).node_neighbor_index is None


def test_truncation_keeps_smallest_indices():
    # 6 quads around node 0 give it 18 distinct neighbours 1..18: the table
    # keeps 1..16; every other node sees 0 and its two element mates
    el = np.array([[0, 1 + 3 * k, 2 + 3 * k, 3 + 3 * k] for k in range(6)], np.int32)
    ang = np.linspace(0, 2 * np.pi, 19)[:-1]
    nodes = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(ang), np.sin(ang)], 1)]).astype(np.float32)
    idx, deg = build_node_neighbor_tables(Mesh(nodes_xy=nodes, elements=el), max_degree=16)
    idx, deg = np.asarray(idx), np.asarray(deg)
    assert deg[0] == 16
    np.testing.assert_array_equal(idx[0], np.arange(1, 17))
    np.testing.assert_array_equal(deg[1:], 3)
    np.testing.assert_array_equal(idx[4, :3], [0, 5, 6])
    idx8, deg8 = build_node_neighbor_tables(Mesh(nodes_xy=nodes, elements=el), max_degree=8)
    np.testing.assert_array_equal(np.asarray(idx8)[0], np.arange(1, 9))
    assert np.asarray(deg8)[0] == 8
