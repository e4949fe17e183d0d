"""Discrepancy map (d2ic_discrepancy_map) vs the oracle's restatement of
discrepancy.py:51-127 on the same pixel assets.

Bars: the map bit-exact (same NaN pattern, same FP32 values: the kernel
rounds every operation separately in the reference's order), the RMS within
rtol 1e-6 (FP64 block partials vs one FP64 sum), and the reference's NaN RMS
when a sample leaves the bilinear domain."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _np_assets(pd):
    return dict(pixel_coords_ref=pd.pixel_coords_ref.detach().cpu().numpy(),
                pixel_nodes=pd.pixel_nodes.detach().cpu().numpy(),
                pixel_shapeN=pd.pixel_shapeN.detach().cpu().numpy(),
                roi_mask_flat=pd.roi_mask_flat.detach().cpu().numpy(), image_shape=tuple(pd.image_shape))


def _smooth_u(nodes, amp, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(size=6)
    x, y = nodes[:, 0] / 100.0, nodes[:, 1] / 100.0
    ux = amp * (c[0] + c[1] * x + np.sin(2.0 * x + c[2] * y))
    uy = amp * (c[3] + c[4] * y + np.cos(1.5 * y + c[5] * x))
    return np.stack([ux, uy], 1).astype(np.float32)


@pytest.mark.parametrize("full_roi", [False, True])
def test_discrepancy_matches_oracle(gpu, full_roi):
    from d2ic_amd import mask_to_mesh_assets, speckle
    from d2ic_amd.discrepancy import compute_discrepancy_map_ref
    from d2ic_amd.pixel_assets import build_pixel_assets

    H, W = 120, 136
    print(f"[case] discrepancy map vs oracle, full_roi={full_roi}", flush=True)
    t = gpu
    m = speckle.rect_mask(H, W, H, W) if full_roi else speckle.rect_mask(H, W, 80, 96)
    mesh, assets = mask_to_mesh_assets(m, 8)
    ref, de = speckle.speckle_pair(H, W, 3, speckle.rigid(0.6, -0.4), noise_sigma=0.5)
    u = _smooth_u(np.asarray(mesh.nodes_xy), 0.8, 4)
    pd = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
    mp, rms = compute_discrepancy_map_ref(ref_image=ref, def_image=de, u_nodal=t.tensor(u, device="cuda"),
                                          pixel_assets=pd)
    mo, rmso = O.discrepancy_map(np.asarray(ref), np.asarray(de), u, _np_assets(pd))
    mp = mp.cpu().numpy()
    n_fin = int(np.isfinite(mo).sum())
    print(f"finite pixels {n_fin}, rms gpu {rms} oracle {rmso}")
    assert n_fin > 1000
    np.testing.assert_array_equal(mp, mo)
    if full_roi:
        # pixels of the last image column/row sample outside: the reference's RMS is NaN
        assert np.isnan(rmso) and np.isnan(rms)
    else:
        assert np.isfinite(rmso)
        np.testing.assert_allclose(rms, rmso, rtol=1e-6)


def test_discrepancy_duplicates_and_empty(gpu):
    import dataclasses

    from d2ic_amd import mask_to_mesh_assets, speckle
    from d2ic_amd.discrepancy import compute_discrepancy_map_ref
    from d2ic_amd.pixel_assets import build_pixel_assets

    print("[case] discrepancy map: duplicated image pixels, empty pixel set", flush=True)
    t = gpu
    H = W = 64
    m = speckle.rect_mask(H, W, 40, 40)
    mesh, assets = mask_to_mesh_assets(m, 8)
    ref, de = speckle.speckle_pair(H, W, 8, speckle.rigid(0.3, 0.2), noise_sigma=0.5)
    u = _smooth_u(np.asarray(mesh.nodes_xy), 0.5, 9)
    pd = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
    # every 5th pixel reports the image pixel of its predecessor
    rf = pd.roi_mask_flat.clone()
    rf[5::5] = rf[4:-1:5][: rf[5::5].numel()]
    pd2 = dataclasses.replace(pd, roi_mask_flat=rf)
    mp, rms = compute_discrepancy_map_ref(ref_image=ref, def_image=de, u_nodal=t.tensor(u, device="cuda"),
                                          pixel_assets=pd2)
    mo, rmso = O.discrepancy_map(np.asarray(ref), np.asarray(de), u, _np_assets(pd2))
    np.testing.assert_array_equal(mp.cpu().numpy(), mo)
    np.testing.assert_allclose(rms, rmso, rtol=1e-6)
    n = pd.pixel_coords_ref.shape[0]
    pd0 = dataclasses.replace(pd, pixel_coords_ref=pd.pixel_coords_ref[:0], pixel_nodes=pd.pixel_nodes[:0],
                              pixel_shapeN=pd.pixel_shapeN[:0], roi_mask_flat=pd.roi_mask_flat[:0])
    assert n > 0
    me, rmse = compute_discrepancy_map_ref(ref_image=ref, def_image=de, u_nodal=t.tensor(u, device="cuda"),
                                           pixel_assets=pd0)
    assert np.all(np.isnan(me.cpu().numpy())) and rmse == 0.0


def test_dic_run_discrepancy_output(gpu):
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets, speckle

    print("[case] DICMeshBased with compute_discrepancy_map", flush=True)
    H = W = 96
    m = speckle.rect_mask(H, W, 64, 64)
    mesh, assets = mask_to_mesh_assets(m, 16)
    ref, de = speckle.speckle_pair(H, W, 2, speckle.rigid(0.4, -0.3), noise_sigma=0.5)
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation="cubic"),
                       MeshDICConfig(max_iters=50, tol=1e-4, reg_strength=0.1, compute_discrepancy_map=True))
    dic.prepare(ref, assets)
    res = dic.run(de)
    u = res.u_nodal.detach().cpu().numpy()
    pd = dic.state.assets.pixel_data
    mo, rmso = O.discrepancy_map(np.asarray(ref), np.asarray(de), u, _np_assets(pd))
    np.testing.assert_array_equal(res.pixel_maps["discrepancy_ref"].cpu().numpy(), mo)
    np.testing.assert_allclose(res.diagnostics.info["discrepancy_rms"], rmso, rtol=1e-6)
    info = res.diagnostics.info
    assert info["discrepancy"] == "discrepancy_ref"
    # the rest of the reference's diagnostics (dic_mesh_based.py:100-155)
    assert info["stage"] == "mesh_based" and info["strain"] == "green_lagrange"
    assert isinstance(info["n_iters"], int) and 0 < info["n_iters"] <= 50
    assert info["history"] == "attached"  # save_history defaults to True
    assert tuple(res.strain.shape) == (mesh.nodes_xy.shape[0], 3)
    # the solved displacement explains most of the residual at u = 0
    _, rms0 = O.discrepancy_map(np.asarray(ref), np.asarray(de), np.zeros_like(u), _np_assets(pd))
    print(f"rms at the solution {rmso:.3f}, at u = 0 {rms0:.3f}")
    assert rmso < 0.5 * rms0
