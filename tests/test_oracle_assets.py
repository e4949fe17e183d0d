"""Oracle pixel assets pinned by the reference's integer fixture, brute force and closed forms."""
from __future__ import annotations

import numpy as np

from d2ic_amd.mask2mesh import mask_to_mesh
from d2ic_amd.pixel_assets import build_node_neighbor_dense
from oracle import oracle as O
from tests._cases import distorted_mesh


def test_reference_fixture_unit_quad():
    """tests/test_pixel_assets.py:6-36 of the reference (chunking is irrelevant here)."""
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    elements = np.array([[0, 1, 2, 3]], np.int32)
    px = np.array([[0.5, 0.5], [0.1, 0.9], [1.5, 1.5]])
    for use_bins in (False, True):
        elts, pn = O.pixel_to_element(px, nodes, elements, use_bins=use_bins)
        assert elts.tolist() == [0, 0, -1]
        assert pn[0].tolist() == [0, 1, 2, 3]
        assert pn[2].tolist() == [0, 0, 0, 0]


def test_binned_search_equals_brute_force():
    rng = np.random.default_rng(0)
    mesh = distorted_mesh(nx=9, ny=8, h=11.0, amp=3.0, seed=5)
    pts = rng.uniform(0, 110, size=(5000, 2)).astype(np.float32)
    pts[:200] = np.round(pts[:200] * 2) / 2  # half-integers hit edges/vertices
    a, _ = O.pixel_to_element(pts, mesh.nodes_xy, mesh.elements, use_bins=False)
    b, _ = O.pixel_to_element(pts, mesh.nodes_xy, mesh.elements, use_bins=True)
    assert np.array_equal(a, b)


def test_first_match_on_shared_edges_is_lowest_index():
    mesh = mask_to_mesh(np.ones((64, 64), bool), 16)
    # pixel centres never sit on grid lines (x.5 vs multiples of 16), so use explicit points
    pts = np.array([[16.0, 8.0], [16.0, 16.0], [48.0, 63.0]], np.float32)
    e, _ = O.pixel_to_element(pts, mesh.nodes_xy, mesh.elements, use_bins=False)
    for p, ei in zip(pts, e):
        cands = [k for k in range(mesh.elements.shape[0])
                 if O.pixel_to_element(p[None], mesh.nodes_xy, mesh.elements[k:k + 1], use_bins=False)[0][0] == 0]
        assert ei == min(cands)


def test_structured_closed_form():
    """On a 16-px structured grid: element = cell index, xi = 2(x - xc)/16 exactly,
    node->pixel lists = ascending pixel ids, degree sum = 4 Np."""
    H = W = 96
    mesh = mask_to_mesh(np.ones((H, W), bool), 16)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W))
    xy = a["pixel_coords_ref"].astype(np.float64)
    gx = np.unique(mesh.nodes_xy[:, 0])
    nx = gx.size
    cx = np.clip(np.searchsorted(gx, xy[:, 0], side="right") - 1, 0, nx - 2)
    cy = np.clip(np.searchsorted(gx, xy[:, 1], side="right") - 1, 0, nx - 2)
    el = mesh.elements
    cell_of = {(int(el[k, 0] % nx), int(el[k, 0] // nx)): k for k in range(el.shape[0])}
    expect = np.array([cell_of[(i, j)] for i, j in zip(cx, cy)])
    interior = xy[:, 0] < gx[-2]
    interior &= xy[:, 1] < gx[-2]
    assert np.array_equal(a["pixel_elts"][interior], expect[interior])
    # exact dyadic shape functions in interior (16 px) cells
    x0 = gx[cx]
    y0 = gx[cy]
    xi = 2 * (xy[:, 0] - (x0 + 8)) / 16
    eta = 2 * (xy[:, 1] - (y0 + 8)) / 16
    N = 0.25 * np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], 1)
    assert np.array_equal(a["pixel_shapeN"][interior], N[interior].astype(np.float32))
    deg = a["node_pixel_degree"]
    assert deg.sum() == 4 * xy.shape[0]
    idx = a["node_pixel_index"]
    for n in range(idx.shape[0]):
        row = idx[n, :deg[n]]
        assert np.all(np.diff(row) > 0)
        assert np.all(idx[n, deg[n]:] == -1)


def test_shape_functions_partition_and_inverse_map():
    mesh = distorted_mesh()
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (80, 90))
    N = a["pixel_shapeN"].astype(np.float64)
    np.testing.assert_allclose(N.sum(1), 1.0, atol=3e-7)
    assert N.min() >= -1e-6
    X = mesh.nodes_xy[a["pixel_nodes"]].astype(np.float64)  # (Np,4,2)
    xmap = np.einsum("pa,pac->pc", N, X)
    np.testing.assert_allclose(xmap, a["pixel_coords_ref"], atol=2e-4)


def test_node_pixel_table_matches_numpy_stable_argsort():
    mesh = distorted_mesh(seed=9)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (80, 90))
    flat = a["pixel_nodes"].reshape(-1)
    order = np.argsort(flat, kind="stable")
    nn = mesh.nodes_xy.shape[0]
    starts = np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=nn))[:-1]])
    for n in range(nn):
        ent = order[starts[n]:starts[n] + a["node_pixel_degree"][n]]
        assert np.array_equal(a["node_pixel_index"][n, :ent.size], ent // 4)
        assert np.array_equal(a["node_N_weight"][n, :ent.size], a["pixel_shapeN"].reshape(-1)[ent])


def test_neighbor_table_oracle_vs_host_numpy_and_weights():
    mesh = mask_to_mesh(np.ones((64, 80), bool), 16)
    i1, d1, w1 = O.node_neighbor_dense(mesh.elements, mesh.nodes_xy, mesh.nodes_xy.shape[0])
    i2, d2, w2 = build_node_neighbor_dense(mesh.elements, mesh.nodes_xy, mesh.nodes_xy.shape[0])
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2) and np.array_equal(w1, w2)
    vals = set(np.unique(w1[w1 > 0]).tolist())
    assert np.float32(1.0 / (np.float32(16.0) + np.float32(1e-6))) in vals
    diag = np.float32(np.sqrt(np.float32(512.0)))
    assert np.float32(1.0 / (diag + np.float32(1e-6))) in vals
    assert d1.max() == 8


def test_roi_filter_counts_match_configs():
    """Np for the 512^2 workload with and without roi_mask (pixel_assets.py:62-79)."""
    from d2ic_amd import speckle

    m = speckle.rect_mask(512, 512, 400, 400)
    mesh = mask_to_mesh(m, 16)
    assert mesh.nodes_xy.shape[0] == 1089 and mesh.elements.shape[0] == 625
    c, *_ = O.pixel_candidates(mesh.nodes_xy, mesh.elements, 512, 512)
    assert c.shape[0] == 160000
    c2, *_ = O.pixel_candidates(mesh.nodes_xy, mesh.elements, 512, 512, roi_mask=m)
    assert c2.shape[0] == 153664
