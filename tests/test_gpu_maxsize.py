"""The largest configuration of the batch and serving mixes on one problem:
4096^2 image, 3280^2 rectangle ROI aligned to the 8 px grid, Q4-8 mesh
(263,169 nodes, 168,100 elements, 10.76 Mpx).  Checked through properties that
do not depend on size: the pixel assets against their closed forms on a
structured grid (pixel_assets.py:21-132, 240-327, 459-498: element = grid
cell, bilinear shape functions, node degree 4 x 64 inside, roi_mask_flat),
the first 8 NLCG iterations against the FP64 oracle at full size (the alpha
sequence exact, J and |grad| histories, iterates within 2e-5 px), and the
discrepancy map bit-exact against the oracle at full size."""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

H = W = 4096
E = 8
SIDE = 3280
SHIFT = (0.35, -0.25)


def test_max_size_assets_solve_and_discrepancy(gpu):
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, speckle
    from d2ic_amd.discrepancy import compute_discrepancy_map_ref
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.mesh_assets import make_mesh_assets
    from d2ic_amd.pixel_assets import build_pixel_assets

    t = gpu
    print("[case] 4096^2 Q4-8 rectangle: assets", flush=True)
    mask = speckle.rect_mask(H, W, SIDE, SIDE)
    mesh = mask_to_mesh(mask, E)
    nodes = np.asarray(mesh.nodes_xy)
    ne = mesh.elements.shape[0]
    assert nodes.shape[0] == 263169 and ne == (SIDE // E) ** 2
    ref, frames = speckle.speckle_sequence_device(H, W, 21, [speckle.rigid(*SHIFT)])
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=mask)
    xy = pa.pixel_coords_ref.cpu().numpy()
    pn = pa.pixel_nodes.cpu().numpy()
    N = pa.pixel_shapeN.cpu().numpy()
    rf = pa.roi_mask_flat.cpu().numpy()
    npx = xy.shape[0]
    assert npx == 64 * ne == 10_758_400
    # pixel centres of the ROI in row-major order
    r0 = (H - SIDE) // 2
    i, j = np.divmod(np.arange(npx, dtype=np.int64), SIDE)
    np.testing.assert_array_equal(xy[:, 0], (j + r0).astype(np.float32) + np.float32(0.5))
    np.testing.assert_array_equal(xy[:, 1], (i + r0).astype(np.float32) + np.float32(0.5))
    np.testing.assert_array_equal(rf, (i + r0) * W + (j + r0))
    # element = grid cell: node 0 is the cell's corner n0 = iy * nx + ix, then n0+1, n0+nx+1, n0+nx
    nx = 513
    cx, cy = (j + r0) // E, (i + r0) // E
    n0 = cy * nx + cx
    np.testing.assert_array_equal(pn, np.stack([n0, n0 + 1, n0 + nx + 1, n0 + nx], 1).astype(np.int32))
    # bilinear shape functions of the reference coordinates xi = 2 (x - xc) / E
    xi = (2.0 * (xy[:, 0] - (cx * E + E / 2)) / E).astype(np.float64)
    eta = (2.0 * (xy[:, 1] - (cy * E + E / 2)) / E).astype(np.float64)
    Nc = 0.25 * np.stack([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)], 1)
    assert np.abs(N - Nc).max() < 1e-6
    deg = pa.node_pixel_degree.cpu().numpy()
    assert int(deg.sum()) == 4 * npx and int(deg.max()) == 4 * E * E
    assert int(pa.node_pixel_index.shape[1]) == 4 * E * E

    print("[case] 4096^2 Q4-8 rectangle: first 8 CG iterations vs the FP64 oracle", flush=True)
    # (the reference's NLCG scales the gradient by 1/Np, so from u = 0 a
    # 10.76 Mpx problem moves ~1e-3 px per iteration: 200 iterations do not
    # reach a 0.35 px shift, for the reference as here; the full-size check is
    # the iterate-by-iterate agreement with the oracle)
    assets = replace(make_mesh_assets(mesh, with_neighbors=True), pixel_data=pa)
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation="cubic"),
                       MeshDICConfig(max_iters=8, tol=1e-9, reg_strength=0.1, save_history=True))
    dic.prepare(ref, assets)
    res = dic.run(frames[0])
    u = res.u_nodal.detach().cpu().numpy()
    hg = res.history.detach().cpu().numpy()
    oa = dict(pixel_coords_ref=xy, pixel_nodes=pn, pixel_shapeN=N, node_pixel_index=pa.node_pixel_index.cpu().numpy(),
              node_N_weight=pa.node_N_weight.cpu().numpy(), node_pixel_degree=deg,
              node_neighbor_index=pa.node_neighbor_index.cpu().numpy(),
              node_neighbor_degree=pa.node_neighbor_degree.cpu().numpy(),
              node_neighbor_weight=pa.node_neighbor_weight.cpu().numpy(),
              node_reg_weight=pa.node_reg_weight.cpu().numpy())
    u0 = np.zeros((nodes.shape[0], 2), np.float32)
    u64, h64, n64 = O.cg_solve(u0, ref.cpu().numpy(), frames[0].cpu().numpy(), oa, alpha_reg=0.1, max_iter=8, tol=1e-9,
                               interp="cubic", dtype=np.float64)
    print(f"J {hg[0, 0]:.5f} -> {hg[-1, 0]:.5f}, alpha {hg[:, 2].tolist()}, max |u - u64| {np.abs(u - u64).max():.2e}")
    assert res.diagnostics.info["n_iters"] == n64 == 8
    np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
    np.testing.assert_allclose(hg[:, 1], h64[:, 1], rtol=2e-4)
    np.testing.assert_allclose(u, u64, atol=2e-5)
    assert hg[-1, 0] < hg[0, 0]

    print("[case] 4096^2 Q4-8 rectangle: discrepancy map vs oracle", flush=True)
    mp, rms = compute_discrepancy_map_ref(ref_image=ref, def_image=frames[0], u_nodal=res.u_nodal, pixel_assets=pa)
    a = dict(pixel_coords_ref=xy, pixel_nodes=pn, pixel_shapeN=N, roi_mask_flat=rf, image_shape=(H, W))
    mo, rmso = O.discrepancy_map(ref.cpu().numpy(), frames[0].cpu().numpy(), u, a)
    np.testing.assert_array_equal(mp.cpu().numpy(), mo)
    np.testing.assert_allclose(rms, rmso, rtol=1e-6)
