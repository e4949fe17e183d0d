"""A whole solve captured into a HIP graph (``SolveGraph``) and replayed per
frame: every replay, with new frames and warm starts written into the
captured buffers and eager GPU work between replays, is bitwise the eager
``d2ic_cg_solve_batch`` solve of the same inputs (u, n_iters, pass count),
for one problem (single-problem kernel) and for three (batch kernel)."""
from __future__ import annotations

import ctypes

import pytest

from tests._cases import structured_case

pytestmark = pytest.mark.gpu


def _problems(specs):
    import torch

    from d2ic_amd import speckle
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan

    out = []
    for (H, side, e, seed, hole) in specs:
        m, mesh, ref, _ = structured_case(H=H, W=H, side=side, e=e, seed=seed, hole=hole)
        pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
        pl = CGPlan(pa)
        pl.set_reference(torch.as_tensor(ref, device="cuda"), "cubic")
        frames = []
        for k, (ux, uy) in enumerate([(0.3, -0.2), (0.55, -0.35), (0.8, -0.5), (1.1, -0.7)]):
            _, de = speckle.speckle_pair(H, H, seed, speckle.rigid(ux, uy), noise_seed=100 + k, noise_sigma=1.0)
            frames.append(torch.as_tensor(de, device="cuda"))
        out.append((pl, frames))
    return out


def _eager(plans, images, us, reg, max_iter, tol):
    import torch

    from d2ic_amd import _lib

    P = len(plans)
    vp = ctypes.c_void_p
    arr = lambda xs: (vp * P)(*[vp(x) for x in xs])  # noqa: E731
    us = [u.clone() for u in us]
    cnts = [torch.zeros(2, dtype=torch.int32, device="cuda") for _ in plans]
    _lib.call("d2ic_cg_solve_batch", P, arr([pl.handle.value for pl in plans]), arr([g.data_ptr() for g in images]),
              (ctypes.c_int32 * P)(*[g.shape[0] for g in images]), (ctypes.c_int32 * P)(*[g.shape[1] for g in images]),
              arr([u.data_ptr() for u in us]), (ctypes.c_float * P)(*([reg] * P)),
              (ctypes.c_int32 * P)(*([max_iter] * P)), (ctypes.c_float * P)(*([tol] * P)), None,
              arr([c.data_ptr() for c in cnts]), arr([c.data_ptr() + 4 for c in cnts]), _lib.stream())
    torch.cuda.synchronize()
    return [(u, int(c[0]), int(c[1])) for u, c in zip(us, (c.cpu() for c in cnts))]


@pytest.mark.parametrize("specs", [
    [(128, 96, 16, 1, 0.0)],
    [(128, 96, 16, 1, 0.0), (160, 128, 8, 2, 20.0), (160, 128, 32, 4, 0.0)],
], ids=["one_problem", "three_problems"])
def test_graph_replays_equal_eager(gpu, specs):
    import torch

    from d2ic_amd.solver_global_cg import SolveGraph

    print(f"[case] solve graph, {len(specs)} problem(s)", flush=True)
    reg, max_iter, tol = 0.1, 60, 1e-4
    probs = _problems(specs)
    plans = [pl for pl, _ in probs]
    g = SolveGraph(plans, [fr[0] for _, fr in probs], reg_strength=reg, max_iter=max_iter, tol=tol)
    warm = [torch.zeros((pl.n_nodes, 2), device="cuda") for pl in plans]
    for k in range(4):
        imgs = [fr[k] for _, fr in probs]
        for dst, src in zip(g.images, imgs):
            dst.copy_(src)
        for dst, src in zip(g.u, warm):
            dst.copy_(src)
        g.replay()
        got = g.results()
        scratch = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        ref = _eager(plans, imgs, warm, reg, max_iter, tol)
        for p, ((ug, ig, pg), (ue, ie, pe)) in enumerate(zip(got, ref)):
            assert (ig, pg) == (ie, pe), (k, p)
            assert ie > 0
            assert torch.equal(ug, ue), (k, p, float((ug - ue).abs().max()))
        warm = [u.clone() for u, _, _ in ref]  # warm start of the next frame
        assert bool(torch.isfinite(scratch))
