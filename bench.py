"""Benchmark: frames/s of the global FE-DIC hot path on one MI355X.

Workload (the 4 Mpx plain-GPU configuration): 2048x2048 synthetic speckle
(spot generator seed 0), ROI = centred 1800x1800 square minus a centred disc of
radius 300 px, structured Q4 mesh with 16 px elements (16,641 nodes, 11,430
elements; roi_mask passed, 2,923,392 pixels).  Frame t displacement
u(x,y,t) = t [eps (x - x0), -0.3 eps (y - y0)] + 1.5 px sin(2 pi x / W) t / 20 on
u_y, eps = 1e-3, plus N(0,1) grey noise.  Each step solves one frame
(GlobalCGSolver: cubic, max_iters=200, tol=1e-3, reg_strength=0.1) warm-started
from the previous frame's solution, then computes nodal Green-Lagrange strain
(DICMeshBased.run).  Images are resident in HBM before the timed region.

Prints ONE JSON line (see the keys below); --gpus > 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stages_leg(mesh, assets, mask, ref, de, reps=5, win=41, search=4):
    """Stages around the solve on the benchmark workload, each timed as the
    median of ``reps`` calls after one untimed call (device-synchronised wall
    time): pixel assets from the mesh (device kernels plus the host spring
    graph), ZNCC at every element centre, IDW of the centre shifts to the
    nodes, Green-Lagrange LSQ strain at the nodes.  ZNCC is reported as window
    pixels x candidates per second and as the share of the chip's VALU issue
    rate it implies: k_zncc issues ~22 VALU instructions per window pixel per
    candidate in its three passes (mean, variance, and the correlation with
    the reference's per-pixel divide; FP64 accumulation), counted in its
    gfx950 assembly; issue rate = 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per
    wave64 instruction x 64 lanes."""
    import torch

    from d2ic_amd.dic_init_motion import project_centers_to_nodes_idw
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_translation_zncc import match_centers
    from d2ic_amd.strain import strain_voigt

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3, out

    c = np.asarray(assets.element_centers_xy, dtype=np.float32)
    ms_pa, pa = timed(lambda: build_pixel_assets(mesh, ref, 1.0, roi_mask=mask))
    ms_zn, (u_c, _) = timed(lambda: match_centers(c, ref, de, win=win, search=search))
    ms_idw, u_n = timed(lambda: project_centers_to_nodes_idw(mesh.nodes_xy, c, u_c))
    nodes = torch.as_tensor(np.asarray(mesh.nodes_xy), dtype=torch.float32, device="cuda")
    ms_st, _ = timed(lambda: strain_voigt(u_n, nodes, assets.node_neighbor_index, assets.node_neighbor_degree))
    zn_pc = float(win * win) * (2 * search + 1) ** 2 * len(c)
    valu_lane_ops_peak = 256 * 4 * 2.4e9 / 4 * 64
    zn_rate = zn_pc / (ms_zn * 1e-3)
    return dict(pixel_assets_ms=ms_pa, n_pixels=int(pa.pixel_coords_ref.shape[0]),
                zncc_ms=ms_zn, zncc_centres=int(len(c)), zncc_win=win, zncc_search=search,
                zncc_pixel_candidates_per_s=zn_rate, zncc_valu_per_pixel_candidate=22,
                zncc_valu_issue_frac=zn_rate * 22 / valu_lane_ops_peak,
                idw_ms=ms_idw, strain_ms=ms_st, n_nodes=int(nodes.shape[0]))


def _oracle_accuracy(p, u_gpu, reg, max_iters, tol, interp="cubic"):
    """max |u - u_oracle| on the nodes that pixels touch, for one problem of the
    batch/serve mixes: the GPU result against the FP32 and FP64 oracle runs of
    the same solve from the same initial guess."""
    from oracle import oracle as O

    ref = p["ref"].cpu().numpy()
    img = p["img"].cpu().numpy()
    oa = O.pixel_assets(p["mesh"].nodes_xy, p["mesh"].elements, ref.shape, roi_mask=p["mask"])
    used = oa["node_pixel_degree"] > 0
    u0 = p["u0"].cpu().numpy()
    out = {}
    for name, dt in (("fp32", np.float32), ("fp64", np.float64)):
        u, _, n = O.cg_solve(u0, ref, img, oa, alpha_reg=reg, max_iter=max_iters, tol=tol, interp=interp, dtype=dt)
        out[f"max_abs_du_vs_oracle_{name}_px"] = float(np.abs(u_gpu - u)[used].max())
        out[f"oracle_{name}_iters"] = int(n)
    return out


def batch_leg(seed=7, n=64, max_iters=200, tol=1e-3, reg=0.1, accuracy_samples=3):
    """64 independent frame-pair solves of uneven size: one persistent launch
    (d2ic_cg_solve_batch) vs one launch per problem (d2ic_cg_solve), plus the
    accuracy of a few sampled problems (sides <= 1024) against the oracle."""
    import ctypes

    import torch

    from d2ic_amd import _lib
    from d2ic_amd.workloads import batch_problems

    t0 = time.perf_counter()
    probs = batch_problems(seed=seed, n=n)
    log(f"[batch] {n} problems prepared in {time.perf_counter() - t0:.1f} s; "
        f"{sum(p['np'] for p in probs) / 1e6:.1f} Mpx, {sum(p['nn'] for p in probs)} nodes")
    st = _lib.stream()

    def run_plain():
        us, its = [], []
        for p in probs:
            u = p["u0"].clone()
            cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
            _lib.call("d2ic_cg_solve", p["plan"].handle, _lib.ptr(p["img"]), p["H"], p["H"], _lib.ptr(u), reg,
                      max_iters, tol, None, _lib.ptr(cnt), ctypes.c_void_p(cnt.data_ptr() + 4), st)
            us.append(u)
            its.append(cnt)
        torch.cuda.synchronize()
        return us, [int(c[0]) for c in its], [int(c[1]) for c in its]

    def run_batch():
        us = [p["u0"].clone() for p in probs]
        cnts = [torch.zeros(2, dtype=torch.int32, device="cuda") for _ in probs]
        P = len(probs)
        vp = ctypes.c_void_p
        arr = lambda xs: (vp * P)(*[vp(x) for x in xs])  # noqa: E731
        _lib.call("d2ic_cg_solve_batch", P, arr([p["plan"].handle.value for p in probs]),
                  arr([p["img"].data_ptr() for p in probs]), (ctypes.c_int32 * P)(*[p["H"] for p in probs]),
                  (ctypes.c_int32 * P)(*[p["H"] for p in probs]), arr([u.data_ptr() for u in us]),
                  (ctypes.c_float * P)(*([reg] * P)), (ctypes.c_int32 * P)(*([max_iters] * P)),
                  (ctypes.c_float * P)(*([tol] * P)), None, arr([c.data_ptr() for c in cnts]),
                  arr([c.data_ptr() + 4 for c in cnts]), st)
        torch.cuda.synchronize()
        return us, [int(c[0]) for c in cnts], [int(c[1]) for c in cnts]

    run_batch()  # warm-up (first launch of the batch shape)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ub, itb, pb = run_batch()
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    up, itp, pp = run_plain()
    t_plain = time.perf_counter() - t0
    same = all(torch.equal(a, b) for a, b in zip(ub, up)) and itb == itp
    mpx_pass = sum(p["np"] * q for p, q in zip(probs, pb)) / 1e6
    acc = []
    if accuracy_samples > 0:
        small = [i for i, p in enumerate(probs) if p["H"] <= 1024][:accuracy_samples]
        for i in small:
            a = _oracle_accuracy(probs[i], ub[i].cpu().numpy(), reg, max_iters, tol)
            a.update(problem=i, side=probs[i]["H"], element_px=probs[i]["e"], roi=probs[i]["roi"],
                     gpu_iters=itb[i])
            acc.append(a)
    sizes = [p["H"] for p in probs]
    res = dict(n_problems=n, sizes={k: sizes.count(k) for k in (512, 1024, 2048, 4096)},
               mpx=sum(p["np"] for p in probs) / 1e6, batched_s=t_batch, plain_s=t_plain,
               batched_frames_per_s=n / t_batch, plain_frames_per_s=n / t_plain, speedup=t_plain / t_batch,
               bitwise_equal=bool(same), iterations_mean=float(np.mean(itb)), pixel_passes_total=int(sum(pb)),
               batched_mpx_pass_per_s=mpx_pass / t_batch,
               accuracy=dict(note="batched results are bitwise the plain ones (same accuracy); sampled problems "
                                  "against the oracle from the same initial guess", samples=acc),
               init="ZNCC (win 41, search 4) at <=4096 element centres + IDW, shared by both paths")
    log(f"[batch] batched {t_batch * 1e3:.1f} ms, plain {t_plain * 1e3:.1f} ms, speedup {t_plain / t_batch:.2f}, "
        f"bitwise equal {same}; accuracy {acc}")
    res["graph_replay"] = _graph_replay(probs, itp, reg, max_iters, tol)
    for p in probs:
        p["plan"].close()
    return res


def _graph_replay(probs, it_plain, reg, max_iters, tol):
    """The batch's 512^2 problems solved one by one, eagerly (d2ic_cg_solve)
    and by replaying one captured solve graph per problem (SolveGraph), from
    the same initial guess: the per-solve launch cost a graph removes."""
    import ctypes

    import torch

    from d2ic_amd import _lib
    from d2ic_amd.solver_global_cg import SolveGraph

    idx = [i for i, p in enumerate(probs) if p["H"] == 512]
    if not idx:
        return None
    graphs = [SolveGraph([probs[i]["plan"]], [probs[i]["img"]], reg_strength=reg, max_iter=max_iters, tol=tol,
                         u=[probs[i]["u0"]]) for i in idx]
    st = _lib.stream()

    def eager():
        us = []
        for i in idx:
            p = probs[i]
            u = p["u0"].clone()
            cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
            _lib.call("d2ic_cg_solve", p["plan"].handle, _lib.ptr(p["img"]), p["H"], p["H"], _lib.ptr(u), reg,
                      max_iters, tol, None, _lib.ptr(cnt), ctypes.c_void_p(cnt.data_ptr() + 4), st)
            us.append(u)
        torch.cuda.synchronize()
        return us

    def replay():
        for i, g in zip(idx, graphs):
            g.u[0].copy_(probs[i]["u0"])
            g.replay()
        torch.cuda.synchronize()

    eager()
    replay()
    t0 = time.perf_counter()
    ue = eager()
    t_eager = time.perf_counter() - t0
    t0 = time.perf_counter()
    replay()
    t_graph = time.perf_counter() - t0
    same = all(torch.equal(g.u[0], u) for g, u in zip(graphs, ue)) and \
        all(g.results()[0][1] == it_plain[i] for g, i in zip(graphs, idx))
    log(f"[batch] graph replay of {len(idx)} 512^2 solves: {t_graph * 1e3:.1f} ms vs eager {t_eager * 1e3:.1f} ms, "
        f"bitwise equal {same}")
    return dict(problems=len(idx), eager_ms=t_eager * 1e3, graph_ms=t_graph * 1e3, bitwise_equal=bool(same),
                note="includes copying each warm start into the captured buffer before its replay")


def serving_leg(seed=11, n_streams=16, rate=200.0, seconds=1.0, max_iters=200, tol=1e-3, reg=0.1):
    """16 camera streams, Poisson arrivals (aggregate `rate` frame pairs/s for
    `seconds`), ROI side log-uniform in [256, 4096], Q4 elements of 8..32 px,
    smooth affine+sinusoid motion; served (continuous batching) vs plain FIFO.

    Steady state: the number of requests waiting (arrived, not finished) is
    sampled through the run; when it still grows at the end of the arrival
    window the server is not keeping up, and the sustained completion rate is
    what the leg reports (latency percentiles then measure queue growth)."""
    import ctypes

    import torch

    from d2ic_amd import _lib
    from d2ic_amd.serving import DONE, StreamServer
    from d2ic_amd.solver_global_cg import CGPlan
    from d2ic_amd.workloads import serve_streams

    t0 = time.perf_counter()
    streams, arrivals = serve_streams(seed=seed, n_streams=n_streams, rate=rate, seconds=seconds)
    log(f"[serve] {n_streams} streams prepared in {time.perf_counter() - t0:.1f} s; {len(arrivals)} requests; "
        f"sides {[s['side'] for s in streams]}")

    # served: continuous batching
    srv = StreamServer(slots=min(64, n_streams), hbm_budget_bytes=200e9, idle_ms=2.0, max_queue=1 << 20)
    for k, s_ in enumerate(streams):
        if not srv.add_stream(k, s_["pa"], s_["ref"], reg_strength=reg, max_iter=max_iters, tol=tol, u0=s_["u0"]):
            raise RuntimeError("stream does not fit the HBM budget")
    fidx = [0] * n_streams
    reqs = []
    depth = []  # (time, requests arrived and not finished)
    t_start = time.perf_counter()
    i = 0
    next_sample = 0.0
    while i < len(arrivals) or srv.busy():
        now = time.perf_counter() - t_start
        while i < len(arrivals) and arrivals[i][0] <= now:
            k = arrivals[i][1]
            reqs.append(srv.submit(k, streams[k]["frames"][fidx[k]], t_arrival=t_start + arrivals[i][0]))
            fidx[k] += 1
            i += 1
        srv.pump()
        if now >= next_sample:
            depth.append((now, i - len(srv.done) - len(srv.unsolved)))
            next_sample += 0.05
        if now > 300 + 3 * seconds:
            raise RuntimeError("serving leg timed out")
    bad = [r for r in reqs if r.status != DONE]
    if bad:
        raise RuntimeError(f"{len(bad)} served requests did not finish: {bad[0].status} {bad[0].reason}")
    t_end = max(r.t_done for r in reqs)
    lat_s = np.array([r.t_done - r.t_arrival for r in reqs])
    served_u = [r.u for r in reqs]
    srv.close()
    # queue growth over the arrival window: depth in its last fifth vs its second fifth
    d_arr = np.array([d for t, d in depth if t <= seconds])
    tt = np.array([t for t, d in depth if t <= seconds])
    q2 = float(d_arr[(tt > 0.2 * seconds) & (tt <= 0.4 * seconds)].mean()) if d_arr.size else 0.0
    q5 = float(d_arr[tt > 0.8 * seconds].mean()) if d_arr.size else 0.0
    done_in_window = sum(1 for r in reqs if r.t_done - t_start <= seconds)
    bounded = q5 <= max(2.0 * q2, q2 + n_streams)

    # plain: one solve per launch, first come first served
    plans = []
    for s_ in streams:
        pl = CGPlan(s_["pa"])
        pl.set_reference(s_["ref"], "cubic")
        plans.append(pl)
    uprev = [s_["u0"].clone() for s_ in streams]
    fidx = [0] * n_streams
    lat_p, same = [], True
    st = _lib.stream()
    torch.cuda.synchronize()
    t_start_p = time.perf_counter()
    for j, (ta, k) in enumerate(arrivals):
        while time.perf_counter() - t_start_p < ta:
            pass
        u = uprev[k].clone()
        cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
        img = streams[k]["frames"][fidx[k]]
        _lib.call("d2ic_cg_solve", plans[k].handle, _lib.ptr(img), int(img.shape[0]), int(img.shape[1]),
                  _lib.ptr(u), reg, max_iters, tol, None, _lib.ptr(cnt), None, st)
        torch.cuda.synchronize()
        lat_p.append(time.perf_counter() - (t_start_p + ta))
        same = same and torch.equal(u, served_u[j])  # requests were served in arrival order too
        uprev[k] = u
        fidx[k] += 1
    t_end_p = time.perf_counter()
    lat_p = np.array(lat_p)
    for pl in plans:
        pl.close()
    n = len(arrivals)
    res = dict(streams=n_streams, offered_rate=rate, seconds=seconds, requests=n,
               sides=[s_["side"] for s_ in streams], elements=[s_["e"] for s_ in streams],
               mpx_per_request_mean=float(np.mean([streams[k]["pa"].pixel_coords_ref.shape[0] for _, k in arrivals]))
               / 1e6,
               served_frames_per_s=n / (t_end - t_start), plain_frames_per_s=n / (t_end_p - t_start_p),
               served_completed_in_window_per_s=done_in_window / seconds,
               queue_depth_2nd_fifth=q2, queue_depth_last_fifth=q5, queue_bounded=bool(bounded),
               served_p50_ms=1e3 * float(np.percentile(lat_s, 50)), served_p99_ms=1e3 * float(np.percentile(lat_s, 99)),
               plain_p50_ms=1e3 * float(np.percentile(lat_p, 50)), plain_p99_ms=1e3 * float(np.percentile(lat_p, 99)),
               latency_note=("steady state: percentiles are load latencies" if bounded else
                             "queue grew through the window: the offered rate exceeds the sustained rate, so the "
                             "percentiles measure queue growth, not load latency"),
               hbm_budget_bytes=200e9, slots=min(64, n_streams), bitwise_equal=bool(same),
               accuracy="served results are bitwise the plain per-solve results (same accuracy)",
               init="ZNCC (win 41, search 4) + IDW on each stream's first frame, before the timed run")
    log(f"[serve] served {res['served_frames_per_s']:.1f} fr/s p50 {res['served_p50_ms']:.1f} ms "
        f"p99 {res['served_p99_ms']:.1f} ms (queue {q2:.1f} -> {q5:.1f}, bounded {bounded}) | plain "
        f"{res['plain_frames_per_s']:.1f} fr/s p50 {res['plain_p50_ms']:.1f} ms p99 {res['plain_p99_ms']:.1f} ms")
    return res


def differentiation_leg(seed=3, n_probes=256, max_iters=200, tol=1e-3, reg=0.1):
    """1024x1024 speckle pair, Q4-16 on an 800x800 ROI, config-1 style field
    (frame 10): forward solve, loss = |u - u_target|^2 + mean E_xx over a
    200x200 window, gradients w.r.t. both images and reg_strength through the
    implicit adjoint, and 256 forward-tangent noise probes."""
    import torch

    from d2ic_amd import speckle
    from d2ic_amd.adjoint import Linearization, green_lagrange_torch, solve_differentiable
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.mesh_assets import build_node_neighbor_tables
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan

    H = W = 1024
    mask = speckle.rect_mask(H, W, 800, 800)
    mesh = mask_to_mesh(mask, 16)
    fn = speckle.affine_sinusoid(10, W, H)
    ref, frames = speckle.speckle_sequence_device(H, W, seed, [fn], noise_sigma=1.0)
    g = frames[0]
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=mask)
    plan = CGPlan(pa)
    nb_idx, nb_deg = build_node_neighbor_tables(mesh)
    X = np.asarray(mesh.nodes_xy, np.float64)
    ux, uy = fn(X[:, 0], X[:, 1])
    act = plan.node_active
    win = torch.tensor((np.abs(X[:, 0] - W / 2) <= 100) & (np.abs(X[:, 1] - H / 2) <= 100), device="cuda") & act
    u_target = torch.tensor(np.stack([ux, uy], 1), dtype=torch.float32, device="cuda")
    act_w, win_w = act.to(torch.float32), win.to(torch.float32)

    def forward_backward():
        f = ref.clone().requires_grad_(True)
        gg = g.clone().requires_grad_(True)
        alpha = torch.tensor(reg, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u = solve_differentiable(plan, f, gg, alpha, max_iter=max_iters, tol=tol, cg_rtol=1e-6)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        E = green_lagrange_torch(u, X, nb_idx, nb_deg)
        # masks as weights (boolean indexing would add torch's serial index backward)
        loss = (((u - u_target) ** 2) * act_w[:, None]).sum() + (E[:, 0, 0] * win_w).sum() / win_w.sum()
        t0 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return f, gg, alpha, u, loss, t_fwd, time.perf_counter() - t0

    # the first call pays one-time costs (lazy loading of torch's autograd
    # kernels); the second is timed
    t0 = time.perf_counter()
    forward_backward()
    t_first = time.perf_counter() - t0
    f, gg, alpha, u, loss, t_fwd, t_bwd = forward_backward()
    lin = Linearization(plan, g, u.detach())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    du_sq = torch.zeros_like(u, dtype=torch.float64)
    exx_sq = torch.zeros(plan.n_nodes, dtype=torch.float64, device="cuda")
    for k0 in range(0, n_probes, 64):
        nv = min(64, n_probes - k0)
        noise = torch.randn((nv, H, W), generator=gen, device="cuda")
        du = lin.tangent(noise, reg, "g", rtol=1e-4)
        du_sq += (du.double() ** 2).sum(0)
        # first-order E_xx change along each probe (central difference around u*),
        # all probes of the batch at once
        e1 = green_lagrange_torch(u.detach()[None] + 1e-3 * du, X, nb_idx, nb_deg)[..., 0, 0]
        e0 = green_lagrange_torch(u.detach()[None] - 1e-3 * du, X, nb_idx, nb_deg)[..., 0, 0]
        exx_sq += (((e1 - e0) / 2e-3).double() ** 2).sum(0)
    torch.cuda.synchronize()
    t_probe = time.perf_counter() - t0
    u_std = torch.sqrt(du_sq / n_probes)[act]
    exx_std = torch.sqrt(exx_sq / n_probes)[win]
    res = dict(image=[H, W], n_pixels=plan.n_pixels, n_nodes=plan.n_nodes, loss=float(loss.detach()),
               forward_solve_s=t_fwd, backward_s=t_bwd, first_call_s=t_first, probes=n_probes, probes_s=t_probe,
               grad_norms=dict(ref=float(f.grad.norm()), deformed=float(gg.grad.norm()), reg=float(alpha.grad)),
               u_std_median_px=float(u_std.median()), exx_std_window_median=float(exx_std.median()),
               check="finite differences of FP64 oracle solves in tests/test_gpu_adjoint.py (rel. err ~2e-4)")
    log(f"[diff] forward {t_fwd * 1e3:.1f} ms, backward {t_bwd * 1e3:.1f} ms, {n_probes} probes {t_probe:.2f} s; "
        f"u_std median {res['u_std_median_px']:.2e} px, exx_std median {res['exx_std_window_median']:.2e}")
    plan.close()
    return res


def local_refine_leg(dic, ref, frame, u0, n_pixels, sweeps=3, reps=5):
    """LocalGaussNewtonSolver on the headline workload (the reference
    tutorial's refinement: 3 sweeps, lam 0.1, max_step 0.2, omega 0.5) from the
    global solution.  Algorithmic bytes per sweep are those of the reference's
    layout: pixel pass 56 B/px (coords, 4 node ids, 4 shape values, f, and r,
    dg/dx, dg/dy written) + node pass 20 B per node->pixel entry (index,
    weight, gathered r, dg/dx, dg/dy) with 4 entries per pixel.  The solver
    runs on the element tiles (per-tile node sums), which reads ~36 B/px, so
    this figure can exceed what the kernels move."""
    import torch

    from d2ic_amd import DICMeshBased, LocalGaussNewtonSolver, MeshDICConfig
    from d2ic_amd.workloads import PLATE

    w = PLATE
    solver = LocalGaussNewtonSolver(lam=0.1, max_step=0.2, omega=0.5, interpolation=w["interp"])
    loc = DICMeshBased(dic.mesh, solver, MeshDICConfig(max_iters=sweeps, tol=w["tol"], reg_strength=w["reg"],
                                                       save_history=True))
    loc.prepare(ref, dic.state.assets)
    loc.set_initial_guess(u0)
    sol = solver.solve(loc.state, frame)  # warm-up
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        sol = solver.solve(loc.state, frame)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    hist = sol.history.cpu().numpy()
    step = float((sol.u_nodal - u0).abs().max()) if u0 is not None else None
    per_sweep = float(np.median(ms)) / sweeps
    b = (56 + 4 * 20) * n_pixels
    res = dict(sweeps=sweeps, ms_per_refine=float(np.median(ms)), ms_per_sweep=per_sweep,
               algorithmic_bytes_per_sweep=b, achieved_GBs=b / (per_sweep * 1e-3) / 1e9,
               rms_residual_per_sweep=hist[:, 0].tolist(), rms_update_per_sweep=hist[:, 1].tolist(),
               max_update_px=step)
    log(f"[local] {sweeps} sweeps {res['ms_per_refine']:.2f} ms ({per_sweep:.2f} ms/sweep, "
        f"{res['achieved_GBs']:.0f} GB/s algorithmic); RMS residual {hist[0, 0]:.3f} -> {hist[-1, 0]:.3f}")
    return res


def discrepancy_leg(dic, ref, frame, u, n_pixels, reps=5):
    """compute_discrepancy_map_ref on the headline workload at the solved u
    (one map + RMS, the RMS read back to the host inside the timed call).
    Algorithmic bytes: 44 B per pixel (coords, 4 node ids, 4 shape values,
    image index) + both images read once + the map written (12 B per image
    pixel)."""
    import torch

    from d2ic_amd.discrepancy import compute_discrepancy_map_ref

    pd = dic.state.assets.pixel_data
    H, W = (int(v) for v in pd.image_shape)
    mp, rms = compute_discrepancy_map_ref(ref_image=ref, def_image=frame, u_nodal=u, pixel_assets=pd)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        mp, rms = compute_discrepancy_map_ref(ref_image=ref, def_image=frame, u_nodal=u, pixel_assets=pd)
        ms.append((time.perf_counter() - t0) * 1e3)
    b = 44 * n_pixels + 12 * H * W
    res = dict(ms_per_map=float(np.median(ms)), rms=float(rms), algorithmic_bytes=b,
               achieved_GBs=b / (float(np.median(ms)) * 1e-3) / 1e9)
    log(f"[discrepancy] map + RMS {res['ms_per_map']:.2f} ms ({res['achieved_GBs']:.0f} GB/s algorithmic), "
        f"RMS {rms:.3f}")
    return res


def measured_traffic():
    """HBM bytes per k_cg_persistent launch on this workload, from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary (scaled by a copy kernel), or None."""
    f = ROOT / "profiles" / "k_cg_persistent_plate2048_traffic.md"
    if not f.exists():
        return None
    import re

    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", f.read_text())
    return float(m.group(1)) if m else None


def _cpu_model():
    """`lscpu` model name of the host (or the platform string)."""
    import platform
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or platform.machine()


def rect_leg(reps=5):
    """The 512x512 single-frame workload end to end on the GPU: ZNCC (win 41,
    search 4) at the element centres, IDW to the nodes, then the global solve
    (cubic, 200 iterations max, tol 1e-3, reg 0.1) and strain; median of
    ``reps`` runs after one untimed run.  Returns the timing and the arrays the
    CPU baseline compares against."""
    import torch

    from d2ic_amd import (DICInitMotion, DICMeshBased, GlobalCGSolver, InitMotionConfig, MeshDICConfig,
                          TranslationZNCCSolver, mask_to_mesh_assets)
    from d2ic_amd.workloads import RECT, rect_single

    w = RECT
    ref_np, de_np, mask = rect_single()
    ref = torch.as_tensor(ref_np, device="cuda")
    de = torch.as_tensor(de_np, device="cuda")
    mesh, assets = mask_to_mesh_assets(mask, w["e"])
    init = DICInitMotion(InitMotionConfig(win=w["win"], search=w["search"]),
                         TranslationZNCCSolver(InitMotionConfig(win=w["win"], search=w["search"])))
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation=w["interp"]),
                       MeshDICConfig(max_iters=w["max_iters"], tol=w["tol"], reg_strength=w["reg"]))
    dic.prepare(ref, assets)
    init.prepare(ref, dic.state.assets)

    def once():
        u0 = init.run(de).u_nodal
        dic.set_initial_guess(u0)
        return dic.run(de)

    once()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        res = once()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    dic.solver.close()
    return dict(ms=float(np.median(ms)), n_iters=int(res.diagnostics.info["n_iters"]),
                u=res.u_nodal.cpu().numpy(), strain=res.strain.cpu().numpy(), mesh=mesh, mask=mask,
                ref=ref_np, de=de_np)


def cpu_baseline_and_accuracy(ref_np, frames_np, mesh, mask, gpu_frames, n_timed, rect):
    """CPU baseline and accuracy block.

    * 4 Mpx plate (the headline workload): the FP32 oracle (a CPU restatement
      of the reference algorithm, C + OpenMP on every host core) solves frame 1
      (warm-up, untimed) and then frames 2..1+n_timed warm-started, each timed;
      the baseline is 1 / median frame time.  The FP64 oracle runs the same
      chain for the accuracy block.
    * 512x512 single frame: ZNCC + IDW + global solve in the FP32 oracle, one
      untimed run, then the median of 5.
    * accuracy: max |u_gpu - u_oracle| over the nodes that pixels touch, per
      frame, against the FP32 and FP64 chains, and the strain difference."""
    from oracle import oracle as O

    from d2ic_amd.mesh_assets import build_node_neighbor_tables, compute_element_centers
    from d2ic_amd.workloads import PLATE, RECT

    w = PLATE
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref_np.shape, roi_mask=mask)
    used = a["node_pixel_degree"] > 0
    nb_idx, nb_deg = (np.asarray(x) for x in build_node_neighbor_tables(mesh))
    X = np.asarray(mesh.nodes_xy, np.float32)
    chains = {}
    frame_s = []
    for name, dt in (("fp32", np.float32), ("fp64", np.float64)):
        u = np.zeros((X.shape[0], 2), dt)
        us, its = [], []
        for k in range(n_timed + 1):
            t0 = time.perf_counter()
            u, _, n = O.cg_solve(u, ref_np, frames_np[k], a, alpha_reg=w["reg"], max_iter=w["max_iters"],
                                 tol=w["tol"], interp=w["interp"], dtype=dt)
            if name == "fp32" and k > 0:
                frame_s.append(time.perf_counter() - t0)
            us.append(u.copy())
            its.append(n)
        chains[name] = (us, its)
        log(f"[cpu] {name} chain: iterations {its}")
    per_frame = []
    for k, (ug, sg) in enumerate(gpu_frames[: n_timed + 1]):
        u32, u64 = chains["fp32"][0][k], chains["fp64"][0][k]
        _, _, v32 = O.strain(u32, X, nb_idx, nb_deg)
        per_frame.append(dict(frame=k + 1, max_abs_du_vs_fp32_px=float(np.abs(ug - u32)[used].max()),
                              max_abs_du_vs_fp64_px=float(np.abs(ug - u64)[used].max()),
                              max_abs_du_fp32_vs_fp64_px=float(np.abs(u32 - u64)[used].max()),
                              max_abs_dstrain_vs_fp32=float(np.abs(sg - v32)[used].max())))
    t_med = float(np.median(frame_s))

    # 512x512 single frame, end to end
    r = RECT
    ra = O.pixel_assets(rect["mesh"].nodes_xy, rect["mesh"].elements, rect["ref"].shape, roi_mask=rect["mask"])
    cen = np.asarray(compute_element_centers(rect["mesh"]), np.float32)
    rused = ra["node_pixel_degree"] > 0

    def rect_once(dt):
        uc, sc = O.zncc(cen, rect["ref"], rect["de"], r["win"], r["search"])
        uc[~(sc >= 0.2)] = 0.0  # InitMotionConfig.score_min default (solver_translation_zncc.py:74-77)
        u0 = O.idw(rect["mesh"].nodes_xy, cen, uc)
        return O.cg_solve(u0, rect["ref"], rect["de"], ra, alpha_reg=r["reg"], max_iter=r["max_iters"],
                          tol=r["tol"], interp=r["interp"], dtype=dt)

    rect_once(np.float32)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        ur32, _, nr32 = rect_once(np.float32)
        ts.append(time.perf_counter() - t0)
    ur64, _, nr64 = rect_once(np.float64)
    rect_acc = dict(max_abs_du_vs_fp32_px=float(np.abs(rect["u"] - ur32)[rused].max()),
                    max_abs_du_vs_fp64_px=float(np.abs(rect["u"] - ur64)[rused].max()),
                    max_abs_du_fp32_vs_fp64_px=float(np.abs(ur32 - ur64)[rused].max()),
                    gpu_iters=rect["n_iters"], oracle_fp32_iters=int(nr32), oracle_fp64_iters=int(nr64))
    cores = O.num_threads()
    model = _cpu_model()
    baseline = dict(value=1.0 / t_med, unit="frames/s", cores=cores, kind="port",
                    sample=(f"4 Mpx plate: FP32 oracle (C, OpenMP, {cores} threads, {model}), frame 1 untimed, "
                            f"then median of frames 2..{n_timed + 1} warm-started ({t_med:.2f} s per frame; "
                            f"iterations {chains['fp32'][1]}); asset build not included"),
                    frame_s=frame_s, cpu_model=model, nproc=os.cpu_count(),
                    rect512=dict(frames_per_s=1.0 / float(np.median(ts)), ms=1e3 * float(np.median(ts)),
                                 sample="512x512 single frame: ZNCC + IDW + global solve, FP32 oracle, one untimed "
                                        "run then the median of 5", gpu_ms=rect["ms"],
                                 gpu_vs_cpu=float(np.median(ts)) * 1e3 / rect["ms"]))
    accuracy = dict(plate_frames=per_frame, rect512=rect_acc,
                    bars="nodal u within 1e-4 px of the FP32 oracle; strain within 1e-6; the FP64 error is stated",
                    note="nodes touched by pixels; the GPU chain and each oracle chain warm-start from their own "
                         "previous frame")
    return baseline, accuracy


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-frames", type=int, default=5,
                    help="CPU baseline: frames timed after one untimed frame (also the frames of the accuracy block)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-problem batched leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-streams leg")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the served leg at the offered 200 frame pairs/s")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation leg")
    ap.add_argument("--no-local", action="store_true", help="skip the local Gauss-Newton refinement leg")
    ap.add_argument("--autotune", action="store_true",
                    help="choose the persistent grid size with CGPlan.autotune before the timed frames")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (scales rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path (--gpus must be 1).",
              file=sys.stderr)
        return 2

    import torch

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} MiB 5 times")
        return 0

    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets
    from d2ic_amd.workloads import PLATE, plate_sequence

    torch.cuda.init()
    w = PLATE
    K, Wm = int(args.steps), int(args.warmup)
    n_acc = max(0, min(K - 1, args.cpu_frames)) if not args.no_cpu_baseline else 0
    log(f"[bench] rendering {K} frames of {w['H']}x{w['W']} speckle on the GPU")
    ref, frames, mask = plate_sequence(max(K, 1))
    mesh, assets = mask_to_mesh_assets(mask, w["e"])
    cfg = MeshDICConfig(max_iters=w["max_iters"], tol=w["tol"], reg_strength=w["reg"], save_history=False)
    solver = GlobalCGSolver(interpolation=w["interp"])
    dic = DICMeshBased(mesh, solver, cfg)
    t0 = time.perf_counter()
    dic.prepare(ref, assets)
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    info = solver.plan.info()
    Np, Nn = info["n_pixels"], info["n_nodes"]
    B_pass = 48 * Np + 16 * Nn  # algorithmic bytes of one pixel-pass (reference asset layout, FP32)
    log(f"[bench] prepared in {t_prep:.2f} s: Np={Np} Nn={Nn} tiles={info['n_tiles']}")

    # warmup: untimed solves of frame 1 from zero
    for _ in range(Wm):
        dic.set_initial_guess(None)
        dic.run(frames[0])
    torch.cuda.synchronize()

    tune = None
    if args.autotune:
        tune = solver.autotune_grid(dic.state, frames[0])
        log(f"[bench] autotuned grid: {tune['blocks']} blocks ({tune['ms']})")
        torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    kt, kms, passes, iters, kept = [], [], [], [], []
    u_prev = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        dic.set_initial_guess(u_prev)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        res = dic.run(frames[k])
        ev1.record(stream)
        u_prev = res.u_nodal
        if k <= n_acc:
            kept.append((res.u_nodal.clone(), res.strain.clone()))
        kt.append((ev0, ev1))
        passes.append(solver.last_pass_count)
        kms.append(solver.plan.kernel_ms())
        iters.append(res.diagnostics.info["n_iters"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    solve_ms = [a.elapsed_time(b) for a, b in kt]
    fps = K / elapsed
    tot_bytes = B_pass * float(sum(passes))
    # algorithmic bytes of the persistent CG kernel / its measured launch time
    achieved = tot_bytes / (sum(kms) * 1e-3) / 1e9
    traffic = measured_traffic()
    log(f"[bench] {K} frames in {elapsed:.3f} s -> {fps:.2f} frames/s; iters {iters}; passes {passes}")

    # grid autotuner (CGPlan.autotune) on the headline plan, after the timed
    # frames: short fixed-pass solves of frame 1 on each candidate grid, the
    # work-based default (0) included; the headline used the default unless
    # --autotune was given
    t_tune = solver.autotune_grid(dic.state, frames[0], cache=False)
    grid_tune = {"default_ms": t_tune["default_ms"], "tuned_blocks": t_tune["blocks"],
                 "tuned_ms": t_tune["ms"][t_tune["blocks"]], "ms": {str(k): v for k, v in t_tune["ms"].items()},
                 "solve": "8 iterations from u = 0, tol 0, median of 3 after one warm-up",
                 "note": "results are bitwise the same on every grid (tests/test_gpu_grid.py)"}
    solver.plan.set_grid(tune["blocks"] if tune else 0)
    log(f"[bench] grid autotune: {grid_tune}")
    # tile-shape autotuner (GlobalCGSolver.autotune_tiling) on the same mesh:
    # each candidate tiling is a plan of its own, timed like the grids; the
    # headline plan is restored afterwards
    t_tile = solver.autotune_tiling(dic.state, frames[0], cache=False)
    tiling_tune = {"default_ms": t_tile["default_ms"], "tuned": "{}/{}".format(*t_tile["tiling"]),
                   "tuned_ms": t_tile["ms"]["{}/{}".format(*t_tile["tiling"])], "ms": t_tile["ms"],
                   "keys": "tile_px/strip_w (tiles close at tile_px pixels; elements wider than 24 px are cut "
                           "into strips of at most strip_w columns)",
                   "solve": "8 iterations from u = 0, tol 0, median of 3 after one warm-up, default grid",
                   "note": "tilings group the tile partials differently: results agree to rounding "
                           "(tests/test_gpu_grid.py)"}
    solver.set_tiling(dic.state)
    if tune:
        solver.plan.set_grid(tune["blocks"])
    log(f"[bench] tiling autotune: {tiling_tune}")

    local = None
    if not args.no_local:
        log("[bench] local refinement leg: 3 Gauss-Newton sweeps from the last frame's solution")
        local = local_refine_leg(dic, ref, frames[K - 1] if K > 0 else frames[0], u_prev, Np)
        if u_prev is not None:
            local["discrepancy"] = discrepancy_leg(dic, ref, frames[K - 1], u_prev, Np)

    stages = None
    if not args.no_local:
        log("[bench] stages leg: pixel assets, ZNCC at every element centre, IDW, strain")
        stages = stages_leg(mesh, dic.state.assets, mask, ref, frames[0])
        log(f"[bench] stages: {stages}")

    batch = serving = None
    if not args.no_batch:
        log("[bench] batched leg: 64 uneven problems")
        batch = batch_leg()
    if not args.no_serve:
        log("[bench] served leg: 16 streams, offered 200 frame pairs/s")
        serving = {"offered_200_per_s": serving_leg(rate=200.0, seconds=args.serve_seconds)}
        log("[bench] served leg: 16 streams, offered 50 frame pairs/s")
        serving["offered_50_per_s"] = serving_leg(rate=50.0, seconds=min(4.0, args.serve_seconds))

    diff = None
    if not args.no_diff:
        log("[bench] differentiation leg: 1024x1024")
        diff = differentiation_leg()

    cpu = accuracy = None
    if n_acc > 0:
        log("[bench] 512x512 single frame on the GPU (ZNCC + IDW + solve)")
        rect = rect_leg()
        log(f"[bench] CPU baseline and accuracy: oracle FP32/FP64 chains over {n_acc + 1} frames, 512x512 frame")
        ref_np = ref.cpu().numpy()
        fr_np = [f.cpu().numpy() for f in frames[: n_acc + 1]]
        gpu_frames = [(u.cpu().numpy(), e.cpu().numpy()) for u, e in kept]
        cpu, accuracy = cpu_baseline_and_accuracy(ref_np, fr_np, mesh, mask, gpu_frames, n_acc, rect)
        log(f"[bench] accuracy: {accuracy}")

    line = {
        "metric": "frames/s at 4 Mpx Q4-16px; p50/p99 latency under load vs plain; HBM GB/s",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plate-with-hole 2048x2048 speckle, Q4-16 mesh (16641 nodes), warm-started frame sequence, "
                        "plain one-solve-per-launch path",
            "image": [w["H"], w["W"]], "element_size_px": w["e"], "n_pixels": Np, "n_nodes": Nn,
            "interpolation": w["interp"], "max_iters": w["max_iters"], "tol": w["tol"], "reg_strength": w["reg"],
            "iterations_per_frame": iters, "pixel_passes_per_frame": passes,
            "grid_blocks": tune["blocks"] if tune else "default (work-based, capped at one block per CU)",
            "autotune_ms": {str(k): v for k, v in tune["ms"].items()} if tune else None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_cg_persistent (one launch per frame)",
            "algorithmic_bytes_per_launch": B_pass * float(np.mean(passes)),
            "bytes_per_pass": B_pass,
            "mean_kernel_ms": float(np.mean(kms)),
            "mean_solve_ms": float(np.mean(solve_ms)),
        },
        "cpu_baseline": cpu,
        "grid_autotune": grid_tune,
        "tiling_autotune": tiling_tune,
        "accuracy": accuracy,
        "batch": batch,
        "serving": serving,
        "differentiation": diff,
        "local_refine": local,
        "stages": stages,
    }
    print(json.dumps(line), flush=True)
    # release device objects before interpreter teardown (profilers finalise then)
    solver.close()
    del dic, solver, frames, ref, res, u_prev
    torch.cuda.synchronize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
