"""GPU NLCG solver vs the CPU oracle (FP32 and FP64 instances)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests._cases import structured_case

pytestmark = pytest.mark.gpu


def _solve_gpu(mesh, ref, de, roi, interp, max_iter, tol, reg, u0=None):
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, make_mesh_assets
    from dataclasses import replace

    cfg = MeshDICConfig(max_iters=max_iter, tol=tol, reg_strength=reg, save_history=True)
    assets = make_mesh_assets(mesh)
    assets = replace(assets, roi_mask=roi)
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation=interp), cfg)
    dic.prepare(ref, assets)
    if u0 is not None:
        dic.set_initial_guess(u0)
    res = dic.run(de)
    return dic, res


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_cg_matches_oracle_at_convergence(gpu, interp):
    """Tight tolerance (3e-5), down to the FP32 noise floor of |grad J|.

    Bar: the GPU solution is no further from the FP64 oracle's minimiser than
    the FP32 oracle (the reference's own FP32 evaluation order) is, with a
    floor of 1e-4 px: max|gpu - u64| < max(1e-4, 1.25 max|u32 - u64|).  It also
    reaches an objective (evaluated in FP64) no higher than the FP32 oracle's
    beyond 1e-6 relative."""
    print(f"[case] cg parity interp={interp}", flush=True)
    m, mesh, ref, de = structured_case(H=128, W=128, side=96, e=16)
    dic, res = _solve_gpu(mesh, ref, de, m, interp, max_iter=600, tol=3e-5, reg=0.1)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    u32, h32, n32 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=600, tol=3e-5, interp=interp)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=600, tol=3e-5, interp=interp,
                               dtype=np.float64)
    ug = res.u_nodal.cpu().numpy()
    used = oa["node_pixel_degree"] > 0
    d64 = np.abs(ug - u64)[used].max()
    d32 = np.abs(ug - u32)[used].max()
    d_ref = np.abs(u32 - u64)[used].max()

    def J64(u):
        return O.objective(u, ref, de, oa, alpha_reg=0.1, interp=interp, dtype=np.float64, with_grad=False)

    jg, j32, j64 = J64(ug), J64(u32), J64(u64)
    print("n_iters gpu/f32/f64", res.diagnostics.info["n_iters"], n32, n64)
    print(f"max|gpu-f64|={d64:.2e} max|gpu-f32|={d32:.2e} max|f32-f64|={d_ref:.2e}")
    print(f"J64(gpu)-J64(u64)={jg - j64:.3e} J64(u32)-J64(u64)={j32 - j64:.3e}")
    assert d64 < max(1e-4, 1.25 * d_ref)
    assert jg <= j32 + 1e-6 * abs(j64)
    # orphan nodes keep their initial value
    assert np.all(ug[~used] == 0.0)


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_cg_first_iterations_track_oracle(gpu, interp):
    """Same first iterates as the FP64 oracle: alpha sequence and J/|grad| history."""
    print(f"[case] cg history interp={interp}", flush=True)
    m, mesh, ref, de = structured_case(H=128, W=128, side=96, e=16, seed=1)
    dic, res = _solve_gpu(mesh, ref, de, m, interp, max_iter=8, tol=1e-9, reg=0.1)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=8, tol=1e-9, interp=interp,
                               dtype=np.float64)
    hg = res.history.cpu().numpy()
    assert res.diagnostics.info["n_iters"] == n64 == 8
    np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
    np.testing.assert_allclose(hg[:, 1], h64[:, 1], rtol=2e-4)
    ug = res.u_nodal.cpu().numpy()
    np.testing.assert_allclose(ug, u64, atol=2e-5)


@pytest.mark.parametrize("e", [20, 24])
def test_cg_partly_covered_elements_track_oracle(gpu, e):
    """20 and 24 px elements cut by a hole: their pixel runs span up to the
    element's full height, so the plan closes tiles at a row limit (the staged
    image box must fit the LDS slot).  Same first iterates as the FP64 oracle
    (alpha sequence exact, J rtol 2e-5, u within 2e-5 px)."""
    print(f"[case] cg partly covered elements e={e}", flush=True)
    m, mesh, ref, de = structured_case(H=176, W=176, side=160, e=e, seed=4, hole=37.0)
    dic, res = _solve_gpu(mesh, ref, de, m, "cubic", max_iter=8, tol=1e-9, reg=0.1)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    counts = np.bincount(oa["pixel_elts"], minlength=mesh.elements.shape[0])
    assert np.any((counts > 0) & (counts < e * e)), "the case must hold partly covered elements"
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=8, tol=1e-9, interp="cubic",
                               dtype=np.float64)
    hg = res.history.cpu().numpy()
    assert res.diagnostics.info["n_iters"] == n64 == 8
    np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
    ug = res.u_nodal.cpu().numpy()
    np.testing.assert_allclose(ug, u64, atol=2e-5)


def test_cg_recovers_translation(gpu):
    print("[case] cg rigid translation", flush=True)
    m, mesh, ref, de = structured_case(H=128, W=128, side=96, e=16, ux=0.37, uy=-0.52)
    dic, res = _solve_gpu(mesh, ref, de, m, "cubic", max_iter=200, tol=1e-3, reg=0.1)
    oa = dic.state.assets.pixel_data
    used = (oa.node_pixel_degree > 0).cpu().numpy()
    u = res.u_nodal.cpu().numpy()[used]
    assert abs(u[:, 0].mean() - 0.37) < 0.05 and abs(u[:, 1].mean() + 0.52) < 0.05


def test_max_iter_zero_and_history_padding(gpu):
    print("[case] max_iter=0", flush=True)
    m, mesh, ref, de = structured_case(H=96, W=96, side=64, e=16)
    dic, res = _solve_gpu(mesh, ref, de, m, "linear", max_iter=0, tol=1e-3, reg=0.1)
    assert res.diagnostics.info["n_iters"] == 0
    assert float(res.u_nodal.abs().max()) == 0.0
    dic2, res2 = _solve_gpu(mesh, ref, de, m, "linear", max_iter=5, tol=1e-9, reg=0.1)
    h = res2.history.cpu().numpy()
    assert h.shape == (5, 3) and np.isfinite(h).all()


def test_batch_equals_alone(gpu):
    """Each problem's result is bit-identical alone and inside a batch."""
    import torch
    from d2ic_amd import _lib
    from d2ic_amd.solver_global_cg import CGPlan
    from d2ic_amd.pixel_assets import build_pixel_assets
    import ctypes

    print("[case] batch vs alone", flush=True)
    cases = [structured_case(H=96, W=96, side=64, e=16, seed=s, ux=0.2 * s, uy=-0.1 * s) for s in range(3)]
    cases.append(structured_case(H=128, W=128, side=96, e=8, seed=7, ux=-0.4, uy=0.3))
    cases.append(structured_case(H=160, W=160, side=128, e=32, seed=9, ux=0.3, uy=0.2))  # strip-split tiles
    plans, imgs, us, alone = [], [], [], []
    for (m, mesh, ref, de) in cases:
        pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
        pl = CGPlan(pa)
        pl.set_reference(torch.as_tensor(ref, device="cuda"), "cubic")
        plans.append(pl)
        imgs.append(torch.as_tensor(de, device="cuda"))
    for pl, g in zip(plans, imgs):
        u = torch.zeros((pl.n_nodes, 2), device="cuda")
        cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
        _lib.call("d2ic_cg_solve", pl.handle, _lib.ptr(g), g.shape[0], g.shape[1], _lib.ptr(u), 0.1, 150, 1e-4, None,
                  _lib.ptr(cnt), None, _lib.stream())
        alone.append((u.cpu().numpy(), int(cnt[0])))
    P = len(plans)
    us = [torch.zeros((pl.n_nodes, 2), device="cuda") for pl in plans]
    cnts = [torch.zeros(2, dtype=torch.int32, device="cuda") for _ in plans]
    arr = lambda ty, vals: (ty * P)(*vals)
    vp = ctypes.c_void_p
    _lib.call("d2ic_cg_solve_batch", P, arr(vp, [pl.handle.value for pl in plans]), arr(vp, [g.data_ptr() for g in imgs]),
              arr(ctypes.c_int32, [g.shape[0] for g in imgs]), arr(ctypes.c_int32, [g.shape[1] for g in imgs]),
              arr(vp, [u.data_ptr() for u in us]), arr(ctypes.c_float, [0.1] * P), arr(ctypes.c_int32, [150] * P),
              arr(ctypes.c_float, [1e-4] * P), None, arr(vp, [c.data_ptr() for c in cnts]), None, _lib.stream())
    for k in range(P):
        assert int(cnts[k][0]) == alone[k][1]
        assert np.array_equal(us[k].cpu().numpy(), alone[k][0])


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_cg_non_square_image_tracks_oracle(gpu, interp):
    """An image wider than tall (80 x 176) with the ROI reaching columns past
    H - 1: the reference's bilinear sampler clips x by H - 1 and y by W - 1
    (solver_global_cg.py:495-500), which only shows on non-square images.
    Same bars as the square case: alpha sequence exact, J / |grad| histories
    and iterates against the FP64 oracle."""
    from d2ic_amd import speckle
    from d2ic_amd.mask2mesh import mask_to_mesh

    print(f"[case] cg non-square 80x176 interp={interp}", flush=True)
    H, W = 80, 176
    m = np.zeros((H, W), bool)
    m[8:72, 8:168] = True
    mesh = mask_to_mesh(m, 16)
    ref, de = speckle.speckle_pair(H, W, 4, speckle.rigid(0.3, -0.25), noise_sigma=1.0)
    dic, res = _solve_gpu(mesh, ref, de, m, interp, max_iter=8, tol=1e-9, reg=0.1)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    assert float(np.asarray(oa["pixel_coords_ref"])[:, 0].max()) > H  # samples beyond column H - 1
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=8, tol=1e-9, interp=interp,
                               dtype=np.float64)
    hg = res.history.cpu().numpy()
    assert res.diagnostics.info["n_iters"] == n64 == 8
    np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
    np.testing.assert_allclose(hg[:, 1], h64[:, 1], rtol=2e-4)
    np.testing.assert_allclose(res.u_nodal.cpu().numpy(), u64, atol=2e-5)


def test_cg_empty_pixel_set_is_spring_only(gpu):
    """No pixel passes the ROI filter (pixel_assets.py:62-73 with an all-false
    roi_mask): J is the spring energy alone and J_img = 0 / max(0, 1)
    (solver_global_cg.py:237-258); the iterates from a rough u0 still track the
    FP64 oracle."""
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, make_mesh_assets, speckle
    from d2ic_amd.mask2mesh import mask_to_mesh
    from dataclasses import replace

    print("[case] cg with an empty pixel set", flush=True)
    t = gpu
    H = W = 96
    mesh = mask_to_mesh(speckle.rect_mask(H, W, 64, 64), 16)
    none = np.zeros((H, W), bool)
    ref, de = speckle.speckle_pair(H, W, 3, speckle.rigid(0.3, 0.1), noise_sigma=1.0)
    u0 = np.random.default_rng(0).normal(scale=0.2, size=(mesh.nodes_xy.shape[0], 2)).astype(np.float32)
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation="cubic"),
                       MeshDICConfig(max_iters=8, tol=1e-9, reg_strength=0.1, save_history=True))
    dic.prepare(ref, replace(make_mesh_assets(mesh), roi_mask=none))
    assert int(dic.state.assets.pixel_data.pixel_coords_ref.shape[0]) == 0
    dic.set_initial_guess(t.tensor(u0, device="cuda"))
    res = dic.run(de)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=none)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=0.1, max_iter=8, tol=1e-9, interp="cubic",
                               dtype=np.float64)
    hg = res.history.cpu().numpy()
    assert res.diagnostics.info["n_iters"] == n64 == 8
    np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
    np.testing.assert_allclose(res.u_nodal.cpu().numpy(), u64, atol=2e-5)
    assert np.abs(u64 - u0).max() > 0.05  # the springs moved the nodes
