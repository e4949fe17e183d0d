"""LocalGaussNewtonSolver (d2ic_lgn_solve) vs the oracle's restatement of
solver_local_gn.py:205-414 on the same pixel assets.

Bars: nodal displacements within 2e-5 px after the sweeps and per-sweep
history (RMS residual, RMS update) within rtol 1e-4 (FP32 sums in different
orders; the device samples at the split cell + fraction position).  Cubic
sampling is parity-unpinned against the reference (im_jax is not available);
linear follows the reference's _bilinear_sample."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

H = W = 96


def _case(seed=5):
    from d2ic_amd import mask_to_mesh_assets, speckle

    m = speckle.rect_mask(H, W, 64, 64)
    mesh, assets = mask_to_mesh_assets(m, 16)
    fn = speckle.affine_field([[2e-3, 1e-3], [-1e-3, 3e-3]], [0.35, -0.25], 0.2, 80.0, W / 2, H / 2)
    ref, de = speckle.speckle_pair(H, W, seed, fn, noise_sigma=0.5)
    return m, mesh, assets, np.asarray(ref), np.asarray(de)


@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_local_gn_matches_oracle(gpu, interp):
    from d2ic_amd import DICMeshBased, LocalGaussNewtonSolver, MeshDICConfig

    print(f"[case] local GN vs oracle, {interp}", flush=True)
    t = gpu
    m, mesh, assets, ref, de = _case()
    solver = LocalGaussNewtonSolver(lam=0.1, max_step=0.2, omega=0.5, interpolation=interp)
    dic = DICMeshBased(mesh, solver, MeshDICConfig(max_iters=5, tol=1e-3, reg_strength=0.1, save_history=True))
    dic.prepare(ref, assets)
    nn = mesh.nodes_xy.shape[0]
    rng = np.random.default_rng(1)
    u0 = (np.tile([0.3, -0.2], (nn, 1)) + rng.normal(scale=0.05, size=(nn, 2))).astype(np.float32)
    dic.set_initial_guess(t.tensor(u0, device="cuda"))
    sol = dic.solver.solve(dic.state, de)
    u_gpu = sol.u_nodal.detach().cpu().numpy()
    hist_gpu = sol.history.detach().cpu().numpy()
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=m)
    u_or, hist_or = O.local_gn_sweeps(u0, ref, de, oa, lam=0.1, max_step=0.2, alpha_reg=0.1, omega=0.5, n_sweeps=5,
                                      interp=interp)
    err = np.abs(u_gpu - u_or).max()
    moved = np.abs(u_or - u0).max()
    print(f"max |u_gpu - u_oracle| = {err:.2e} px (update size {moved:.3f} px)")
    print("history gpu", hist_gpu.tolist())
    print("history oracle", hist_or.tolist())
    assert moved > 1e-2
    assert err < 2e-5
    np.testing.assert_allclose(hist_gpu, hist_or, rtol=1e-4, atol=1e-7)
    # the full pipeline (strain, diagnostics) runs on the local solver too
    res = dic.run(de)
    np.testing.assert_allclose(res.u_nodal.detach().cpu().numpy(), u_gpu, atol=0, rtol=0)


def test_local_gn_zero_sweeps_and_errors(gpu):
    from d2ic_amd import DICMeshBased, LocalGaussNewtonSolver, MeshDICConfig

    print("[case] local GN: zero sweeps, compile check", flush=True)
    t = gpu
    m, mesh, assets, ref, de = _case(seed=9)
    with pytest.raises(RuntimeError):
        LocalGaussNewtonSolver().solve(None, de)
    dic = DICMeshBased(mesh, LocalGaussNewtonSolver(), MeshDICConfig(max_iters=0, tol=1e-3, reg_strength=0.1,
                                                                      save_history=True))
    dic.prepare(ref, assets)
    nn = mesh.nodes_xy.shape[0]
    u0 = np.full((nn, 2), 0.25, np.float32)
    dic.set_initial_guess(t.tensor(u0, device="cuda"))
    sol = dic.solver.solve(dic.state, de)
    np.testing.assert_array_equal(sol.u_nodal.cpu().numpy(), u0)
    assert tuple(sol.history.shape) == (0, 2)


@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_tile_layout_matches_table_layout(gpu, interp):
    """d2ic_lgn_solve_plan (per-tile node sums over the solver plan's element
    tiles) against d2ic_lgn_solve (the reference's node->pixel table): same
    sweeps up to FP32 summation order, on a mesh with 32 px elements (tiles
    split into 16-column strips) and one with 8 px elements."""
    from d2ic_amd import DICMeshBased, LocalGaussNewtonSolver, MeshDICConfig, mask_to_mesh_assets, speckle

    t = gpu
    for e in (32, 8):
        print(f"[case] local GN tiles vs table, e={e}, {interp}", flush=True)
        m = speckle.rect_mask(H, W, 64, 64)
        mesh, assets = mask_to_mesh_assets(m, e)
        fn = speckle.affine_field([[2e-3, 1e-3], [-1e-3, 3e-3]], [0.35, -0.25], 0.2, 80.0, W / 2, H / 2)
        ref, de = speckle.speckle_pair(H, W, 11, fn, noise_sigma=0.5)
        nn = mesh.nodes_xy.shape[0]
        u0 = t.tensor(np.tile([0.3, -0.2], (nn, 1)).astype(np.float32), device="cuda")
        out = {}
        for layout in ("tiles", "table"):
            dic = DICMeshBased(mesh, LocalGaussNewtonSolver(lam=0.1, max_step=0.2, omega=0.5, interpolation=interp,
                                                            layout=layout),
                               MeshDICConfig(max_iters=4, tol=1e-3, reg_strength=0.1, save_history=True))
            dic.prepare(ref, assets)
            dic.set_initial_guess(u0.clone())
            sol = dic.solver.solve(dic.state, de)
            out[layout] = (sol.u_nodal.cpu().numpy(), sol.history.cpu().numpy())
        du = np.abs(out["tiles"][0] - out["table"][0]).max()
        moved = np.abs(out["table"][0] - u0.cpu().numpy()).max()
        print(f"e={e}: max |u_tiles - u_table| = {du:.2e} px (update {moved:.3f} px)")
        assert moved > 1e-2
        assert du < 2e-6
        np.testing.assert_allclose(out["tiles"][1], out["table"][1], rtol=1e-5, atol=1e-7)
