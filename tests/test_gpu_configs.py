"""The HIP path on the benchmark configurations at their full sizes, against
the oracle (the FP32 run stands in for the reference's default JAX run, the
FP64 run for its x64 run).

* 512x512, 400x400 ROI, Q4-16, rigid (0.37, -0.52) px: ZNCC (win 41,
  search 4) -> IDW -> global solve (cubic, 200 / 1e-3 / 0.1) end to end.
* 2048x2048 plate with hole, Q4-16 (2,923,392 pixels): the first 3 frames of
  the warm-started sequence through BatchMeshBased, cubic and linear.
* The uneven 64-problem batch (sides 512..4096, elements 8/16/32 px): one
  batched launch bitwise equal to 64 single solves, and sampled problems'
  first 8 iterates against the FP64 oracle.
* The 16-stream serving mix: every served request bitwise equal to the plain
  per-solve path.
* 1024x1024 differentiation: the adjoint's image and reg_strength gradients
  and a noise-probe tangent against the FP64 oracle's finite differences.

Bars: integer ZNCC shifts bit-exact; nodal u within 1e-4 px of
the FP32 oracle; Green-Lagrange strain within 1e-6; the FP64 error is printed
and bounded.  Linear mode at 2048^2 is the exception, with numbers: there the
FP32 and FP64 oracle runs themselves differ by ~1e-3 px (the bilinear
gradient jumps at pixel-cell edges, and FP32 coordinates near x = 2000 are
quantised to 1.2e-4 px, which moves samples across them), so no FP32 order of
evaluation pins the result to 1e-4 px; the GPU must then be no further from
the FP64 run than the FP32 oracle is."""
from __future__ import annotations

import ctypes
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _used(oa):
    return oa["node_pixel_degree"] > 0


def test_config_rect512_zncc_idw_solve(gpu):
    from d2ic_amd import (DICInitMotion, DICMeshBased, GlobalCGSolver, InitMotionConfig, MeshDICConfig,
                          TranslationZNCCSolver, mask_to_mesh_assets)
    from d2ic_amd.mesh_assets import build_node_neighbor_tables, compute_element_centers
    from d2ic_amd.workloads import RECT, rect_single

    print("[case] 512^2 rect: ZNCC -> IDW -> global solve", flush=True)
    t = gpu
    w = RECT
    ref, de, mask = rect_single()
    mesh, assets = mask_to_mesh_assets(mask, w["e"])
    icfg = InitMotionConfig(win=w["win"], search=w["search"])
    init = DICInitMotion(icfg, TranslationZNCCSolver(icfg))
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation=w["interp"]),
                       MeshDICConfig(max_iters=w["max_iters"], tol=w["tol"], reg_strength=w["reg"], save_history=True))
    dic.prepare(ref, assets)
    init.prepare(t.as_tensor(ref, device="cuda"), dic.state.assets)
    zs = init.solver.solve(init._state, de)
    r0 = init.run(de)
    assert r0.diagnostics.info == {"stage": "init_motion", "matches": int((~t.isnan(zs.scores)).sum()),
                                   "note": "IDW projection"}
    cen = np.asarray(compute_element_centers(mesh), np.float32)
    uc, sc = O.zncc(cen, ref, de, w["win"], w["search"])
    uc[~(sc >= icfg.score_min)] = 0.0
    sc[~(sc >= icfg.score_min)] = np.nan
    np.testing.assert_array_equal(zs.u_centers.cpu().numpy(), uc)  # integer shifts bit-exact
    np.testing.assert_allclose(zs.scores.cpu().numpy(), sc, rtol=1e-5, equal_nan=True)
    u0 = O.idw(mesh.nodes_xy, cen, uc)
    np.testing.assert_allclose(r0.u_nodal.cpu().numpy(), u0, rtol=1e-5, atol=1e-6)
    dic.set_initial_guess(r0.u_nodal)
    res = dic.run(de)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=mask)
    used = _used(oa)
    u32, h32, n32 = O.cg_solve(u0, ref, de, oa, alpha_reg=w["reg"], max_iter=w["max_iters"], tol=w["tol"],
                               interp=w["interp"], dtype=np.float32)
    u64, h64, n64 = O.cg_solve(u0, ref, de, oa, alpha_reg=w["reg"], max_iter=w["max_iters"], tol=w["tol"],
                               interp=w["interp"], dtype=np.float64)
    ug = res.u_nodal.cpu().numpy()
    n_gpu = res.diagnostics.info["n_iters"]
    e32, e64 = (float(np.abs(ug - x)[used].max()) for x in (u32, u64))
    print(f"n_iters gpu/fp32/fp64 {n_gpu}/{n32}/{n64}; max|du| vs fp32 {e32:.2e} vs fp64 {e64:.2e} "
          f"(fp32 vs fp64 {np.abs(u32 - u64)[used].max():.2e}); |u - truth| max "
          f"{np.abs(ug - [w['ux'], w['uy']])[used].max():.3f} px")
    assert abs(n_gpu - n32) <= 1
    assert e32 <= 1e-4 and e64 <= 1e-4
    nb_idx, nb_deg = (np.asarray(x) for x in build_node_neighbor_tables(mesh))
    _, _, v32 = O.strain(u32, mesh.nodes_xy, nb_idx, nb_deg)
    ds = float(np.abs(res.strain.cpu().numpy() - v32)[used].max())
    print(f"max|strain - oracle strain(u_fp32)| {ds:.2e}")
    assert ds <= 1e-6


@pytest.mark.timeout(900)
@pytest.mark.parametrize("interp", ["cubic", "linear"])
def test_config_plate2048_sequence(gpu, interp):
    from d2ic_amd import BatchConfig, BatchMeshBased, DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets
    from d2ic_amd.mesh_assets import build_node_neighbor_tables
    from d2ic_amd.workloads import PLATE, plate_sequence

    print(f"[case] 2048^2 plate, 3 warm-started frames, {interp}", flush=True)
    w = PLATE
    n_frames = 3
    ref, frames, mask = plate_sequence(n_frames)
    mesh, assets = mask_to_mesh_assets(mask, w["e"])
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation=interp),
                       MeshDICConfig(max_iters=w["max_iters"], tol=w["tol"], reg_strength=w["reg"], save_history=True))
    runner = BatchMeshBased(ref, assets, dic, BatchConfig())
    runner.before(frames)
    out = runner.sequence(frames)
    assert out.diagnostics.info["n_frames"] == n_frames
    ref_np = ref.cpu().numpy()
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref_np.shape, roi_mask=mask)
    assert oa["pixel_coords_ref"].shape[0] == 2923392
    used = _used(oa)
    nb_idx, nb_deg = (np.asarray(x) for x in build_node_neighbor_tables(mesh))
    X = np.asarray(mesh.nodes_xy, np.float32)
    u32 = np.zeros((X.shape[0], 2), np.float32)
    u64 = np.zeros((X.shape[0], 2), np.float64)
    for k, res in enumerate(out.results):
        t0 = time.perf_counter()
        fr = frames[k].cpu().numpy()
        u32, h32, n32 = O.cg_solve(u32, ref_np, fr, oa, alpha_reg=w["reg"], max_iter=w["max_iters"], tol=w["tol"],
                                   interp=interp, dtype=np.float32)
        u64, h64, n64 = O.cg_solve(u64, ref_np, fr, oa, alpha_reg=w["reg"], max_iter=w["max_iters"], tol=w["tol"],
                                   interp=interp, dtype=np.float64)
        ug = res.u_nodal.cpu().numpy()
        hg = res.history.cpu().numpy()
        n_gpu = res.diagnostics.info["n_iters"]
        e32, e64, d = (float(np.abs(a - b)[used].max()) for a, b in ((ug, u32), (ug, u64), (u32, u64)))
        sg = res.strain.cpu().numpy()
        print(f"frame {k + 1}: n_iters gpu/fp32/fp64 {n_gpu}/{n32}/{n64}; max|du| vs fp32 {e32:.2e} vs fp64 "
              f"{e64:.2e} (fp32 vs fp64 {d:.2e}); oracle {time.perf_counter() - t0:.1f} s", flush=True)
        assert n_gpu == n32 == n64 == w["max_iters"]
        if interp == "cubic":
            np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
            assert e32 <= 1e-4 and e64 <= 1e-4
            _, _, v32 = O.strain(u32, X, nb_idx, nb_deg)
            ds = float(np.abs(sg - v32)[used].max())
            print(f"  max|strain - oracle strain(u_fp32)| {ds:.2e}")
            assert ds <= 1e-6
        else:
            assert e64 <= max(1e-4, 1.25 * d), (e64, d)
            _, _, vg = O.strain(ug, X, nb_idx, nb_deg)  # the strain kernel on the GPU's own u
            assert float(np.abs(sg - vg)[used].max()) <= 1e-6
    dic.solver.close()


def test_config_batch64(gpu):
    from d2ic_amd import _lib
    from d2ic_amd.workloads import batch_problems

    print("[case] 64-problem uneven batch", flush=True)
    t = gpu
    reg, max_iters, tol = 0.1, 200, 1e-3
    probs = batch_problems(seed=7, n=64)
    sizes = sorted({p["H"] for p in probs})
    print(f"sizes {sizes}, elements {sorted({p['e'] for p in probs})}, {sum(p['np'] for p in probs) / 1e6:.1f} Mpx",
          flush=True)
    st = _lib.stream()
    alone = []
    for p in probs:
        u = p["u0"].clone()
        cnt = t.zeros(2, dtype=t.int32, device="cuda")
        _lib.call("d2ic_cg_solve", p["plan"].handle, _lib.ptr(p["img"]), p["H"], p["H"], _lib.ptr(u), reg, max_iters,
                  tol, None, _lib.ptr(cnt), ctypes.c_void_p(cnt.data_ptr() + 4), st)
        alone.append((u, cnt))
    P = len(probs)
    vp = ctypes.c_void_p
    arr = lambda xs: (vp * P)(*[vp(x) for x in xs])  # noqa: E731
    us = [p["u0"].clone() for p in probs]
    cnts = [t.zeros(2, dtype=t.int32, device="cuda") for _ in probs]
    _lib.call("d2ic_cg_solve_batch", P, arr([p["plan"].handle.value for p in probs]),
              arr([p["img"].data_ptr() for p in probs]), (ctypes.c_int32 * P)(*[p["H"] for p in probs]),
              (ctypes.c_int32 * P)(*[p["H"] for p in probs]), arr([u.data_ptr() for u in us]),
              (ctypes.c_float * P)(*([reg] * P)), (ctypes.c_int32 * P)(*([max_iters] * P)),
              (ctypes.c_float * P)(*([tol] * P)), None, arr([c.data_ptr() for c in cnts]),
              arr([c.data_ptr() + 4 for c in cnts]), st)
    t.cuda.synchronize()
    for i in range(P):
        assert t.equal(cnts[i], alone[i][1]), i
        assert t.equal(us[i], alone[i][0]), i
    # sampled problems (sides <= 1024, one per element size where possible): first 8 iterates vs FP64
    picked, seen = [], set()
    for i, p in enumerate(probs):
        if p["H"] <= 1024 and p["e"] not in seen:
            picked.append(i)
            seen.add(p["e"])
    assert len(picked) >= 2
    for i in picked:
        p = probs[i]
        u = p["u0"].clone()
        hist = t.empty((8, 3), dtype=t.float32, device="cuda")
        cnt = t.zeros(2, dtype=t.int32, device="cuda")
        _lib.call("d2ic_cg_solve", p["plan"].handle, _lib.ptr(p["img"]), p["H"], p["H"], _lib.ptr(u), reg, 8, 1e-9,
                  _lib.ptr(hist), _lib.ptr(cnt), None, st)
        ref = p["ref"].cpu().numpy()
        oa = O.pixel_assets(p["mesh"].nodes_xy, p["mesh"].elements, ref.shape, roi_mask=p["mask"])
        u64, h64, n64 = O.cg_solve(p["u0"].cpu().numpy(), ref, p["img"].cpu().numpy(), oa, alpha_reg=reg, max_iter=8,
                                   tol=1e-9, interp="cubic", dtype=np.float64)
        hg = hist.cpu().numpy()
        err = float(np.abs(u.cpu().numpy() - u64)[_used(oa)].max())
        print(f"problem {i} ({p['H']}^2, e={p['e']}, {p['roi']}): 8 iterates, max|du| vs fp64 {err:.2e}")
        assert int(cnt[0]) == n64 == 8
        np.testing.assert_array_equal(hg[:, 2], h64[:, 2].astype(np.float32))
        np.testing.assert_allclose(hg[:, 0], h64[:, 0], rtol=2e-5)
        np.testing.assert_allclose(hg[:, 1], h64[:, 1], rtol=2e-4)
        assert err <= 2e-5
    for p in probs:
        p["plan"].close()


@pytest.mark.timeout(600)
def test_config_serve16_mix(gpu):
    """16 Poisson streams (0.5 s of arrivals at 200 frame pairs/s) through
    StreamServer; each request bitwise equal to the plain per-solve path (frame
    k+1 of a stream starts from frame k's result on both paths)."""
    from d2ic_amd import _lib
    from d2ic_amd.serving import DONE, StreamServer
    from d2ic_amd.solver_global_cg import CGPlan
    from d2ic_amd.workloads import serve_streams

    print("[case] 16-stream serving mix", flush=True)
    t = gpu
    reg, max_iters, tol = 0.1, 200, 1e-3
    streams, arrivals = serve_streams(seed=11, n_streams=16, rate=200.0, seconds=0.5)
    print(f"{len(arrivals)} requests; sides {[s['side'] for s in streams]}", flush=True)
    srv = StreamServer(slots=16, hbm_budget_bytes=200e9, idle_ms=2.0, max_queue=1 << 20)
    try:
        for k, s_ in enumerate(streams):
            assert srv.add_stream(k, s_["pa"], s_["ref"], reg_strength=reg, max_iter=max_iters, tol=tol, u0=s_["u0"])
        fidx = [0] * len(streams)
        reqs = []
        t0 = time.perf_counter()
        i = 0
        while i < len(arrivals) or srv.busy():
            now = time.perf_counter() - t0
            while i < len(arrivals) and arrivals[i][0] <= now:
                k = arrivals[i][1]
                reqs.append(srv.submit(k, streams[k]["frames"][fidx[k]], t_arrival=t0 + arrivals[i][0]))
                fidx[k] += 1
                i += 1
            srv.pump()
            assert time.perf_counter() - t0 < 300
        assert all(r.status == DONE for r in reqs)
    finally:
        srv.close()
    st = _lib.stream()
    uprev = [s_["u0"].clone() for s_ in streams]
    fidx = [0] * len(streams)
    plans = []
    for s_ in streams:
        pl = CGPlan(s_["pa"])
        pl.set_reference(s_["ref"], "cubic")
        plans.append(pl)
    for j, (_, k) in enumerate(arrivals):
        u = uprev[k].clone()
        cnt = t.zeros(2, dtype=t.int32, device="cuda")
        img = streams[k]["frames"][fidx[k]]
        _lib.call("d2ic_cg_solve", plans[k].handle, _lib.ptr(img), int(img.shape[0]), int(img.shape[1]), _lib.ptr(u),
                  reg, max_iters, tol, None, _lib.ptr(cnt), None, st)
        assert int(cnt[0]) == reqs[j].n_iters, j
        assert t.equal(u, reqs[j].u), j
        uprev[k] = u
        fidx[k] += 1
    for pl in plans:
        pl.close()


@pytest.mark.timeout(900)
def test_config_diff1024_adjoint(gpu):
    """1024x1024 pair, Q4-16 on an 800x800 ROI, the plate field at frame 10.
    Loss = |u - u_true|^2 over the moving nodes + mean E_xx over the nodes of a
    200x200 window.  At the FP64 minimiser u* (oracle NLCG + Newton), the HIP
    adjoint's dL/dg and dL/df at 32 random pixels each and dL/d reg_strength
    are checked against the implicit-function derivative built from the FP64
    oracle alone: -lambda . dG/dtheta, with dG/dtheta from central differences
    of the oracle gradient (h = 1e-2 grey per pixel, 1e-4 for reg_strength) and
    lambda = H^-1 dL/du from CG on central-difference Hessian-vector products
    of the same gradient.  One noise probe's forward tangent is checked the
    same way.  Bar: 2e-3 of max |reference derivative| (the reference's JAX
    gradients cannot run here: parity against them is unpinned)."""
    from d2ic_amd import speckle
    from d2ic_amd.adjoint import Linearization, green_lagrange_torch
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.mesh_assets import build_node_neighbor_tables
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan
    from tests._fd import cg_fd, oracle_grad, tight_solution

    print("[case] 1024^2 differentiation: adjoint vs oracle finite differences", flush=True)
    t = gpu
    H = W = 1024
    alpha = 0.1
    mask = speckle.rect_mask(H, W, 800, 800)
    mesh = mask_to_mesh(mask, 16)
    fn = speckle.affine_sinusoid(10, W, H)
    ref_d, frames = speckle.speckle_sequence_device(H, W, 3, [fn], noise_sigma=1.0)
    g_d = frames[0]
    ref, de = ref_d.cpu().numpy(), g_d.cpu().numpy()
    pa = build_pixel_assets(mesh, ref_d, 1.0, roi_mask=mask)
    plan = CGPlan(pa)
    plan.set_reference(ref_d, "cubic")
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=mask)
    act = plan.node_active.cpu().numpy()
    X = np.asarray(mesh.nodes_xy, np.float64)
    ux, uy = fn(X[:, 0], X[:, 1])
    u_true = np.stack([ux, uy], 1) * act[:, None]
    t0 = time.perf_counter()
    u_star = tight_solution(ref, de, oa, u_true, alpha, act[:, None], nlcg_iters=100)
    grad = oracle_grad(ref, de, oa, alpha, act[:, None])
    g_res = float(np.abs(grad(u_star)).max())
    print(f"u* in {time.perf_counter() - t0:.1f} s, max|grad J(u*)| {g_res:.2e}, |u* - u_true| max "
          f"{np.abs(u_star - u_true)[act].max():.3f} px", flush=True)
    assert g_res < 1e-9

    nb_idx, nb_deg = (np.asarray(x) for x in build_node_neighbor_tables(mesh))
    win = act & (np.abs(X[:, 0] - W / 2) <= 100) & (np.abs(X[:, 1] - H / 2) <= 100)
    ut = t.tensor(u_star, dtype=t.float64, device="cuda", requires_grad=True)
    E = green_lagrange_torch(ut, X, nb_idx, nb_deg)
    act_t, win_t = t.tensor(act, device="cuda"), t.tensor(win, device="cuda")
    L = ((ut - t.tensor(u_true, device="cuda"))[act_t] ** 2).sum() + E[win_t, 0, 0].mean()
    L.backward()
    dLdu = ut.grad.cpu().numpy() * act[:, None]

    # HIP adjoint
    lin = Linearization(plan, g_d, u_star.astype(np.float32))
    lam = lin.solve(t.tensor(dLdu[None], dtype=t.float32, device="cuda"), alpha, rtol=1e-7)[0]
    gf, gg = (x.cpu().numpy() for x in lin.image_grads(lam))
    galpha = float(-(lam.double() * lin.spring(t.tensor(u_star[None], dtype=t.float32, device="cuda"))[0].double())
                   .sum())

    # the oracle's implicit-function derivative
    t0 = time.perf_counter()
    lam_o = cg_fd(grad, u_star, dLdu, rtol=1e-10, max_iter=2000)
    print(f"oracle lambda in {time.perf_counter() - t0:.1f} s; |lambda_gpu - lambda_oracle| / |lambda_oracle| "
          f"{np.abs(lam.cpu().numpy() - lam_o).max() / np.abs(lam_o).max():.2e}", flush=True)
    rng = np.random.default_rng(32)
    rows, cols = np.nonzero(mask[140:884, 140:884])
    pick = rng.choice(len(rows), size=32, replace=False)
    h = 1e-2
    for name, got in (("g", gg), ("f", gf)):
        ref_v, adj = [], []
        for k in pick:
            r, c = rows[k] + 140, cols[k] + 140
            Gs = []
            for sgn in (1.0, -1.0):
                rr, dd = ref.copy(), de.copy()
                (dd if name == "g" else rr)[r, c] += np.float32(sgn * h)
                Gs.append(oracle_grad(rr, dd, oa, alpha, act[:, None])(u_star))
            dG = (Gs[0] - Gs[1]) / (2 * h)
            ref_v.append(-float((lam_o * dG).sum()))
            adj.append(float(got[r, c]))
        ref_v, adj = np.array(ref_v), np.array(adj)
        err = np.abs(ref_v - adj).max() / np.abs(ref_v).max()
        print(f"dL/d{name} at 32 pixels: max |reference| {np.abs(ref_v).max():.3e}, rel err {err:.2e}", flush=True)
        assert err < 2e-3
    ha = 1e-4
    dGa = (oracle_grad(ref, de, oa, alpha + ha, act[:, None])(u_star)
           - oracle_grad(ref, de, oa, alpha - ha, act[:, None])(u_star)) / (2 * ha)
    ra = -float((lam_o * dGa).sum())
    print(f"dL/dalpha: reference {ra:.6e} adjoint {galpha:.6e}", flush=True)
    assert abs(ra - galpha) < 2e-3 * abs(ra)

    # one forward-tangent noise probe
    dg = np.random.default_rng(5).normal(size=(H, W)).astype(np.float32)
    du = lin.tangent(t.tensor(dg[None], device="cuda"), alpha, "g", rtol=1e-7)[0].cpu().numpy()
    eps = 1e-2
    rhs = -(oracle_grad(ref, de + eps * dg, oa, alpha, act[:, None])(u_star)
            - oracle_grad(ref, de - eps * dg, oa, alpha, act[:, None])(u_star)) / (2 * eps)
    du_o = cg_fd(grad, u_star, rhs, rtol=1e-10, max_iter=2000)
    err = np.abs(du - du_o)[act].max() / np.abs(du_o[act]).max()
    print(f"noise-probe tangent: rel err {err:.2e}", flush=True)
    assert err < 2e-3
    assert np.all(du[~act] == 0.0)
    plan.close()
