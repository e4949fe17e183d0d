"""d2ic_cg_solve_batch_noalloc (the allocation-free variant of
d2ic_cg_solve_batch) through the C ABI: results bitwise equal to
d2ic_cg_solve_batch on two calls in a row, the second with a smaller input,
so the second call runs on the scratch array the first one left behind; more
than 8 problems are rejected before any device work."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests._cases import structured_case

pytestmark = pytest.mark.gpu


def _plans(specs):
    import torch

    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan

    out = []
    for (H, side, e, seed, ux, uy, hole) in specs:
        m, mesh, ref, de = structured_case(H=H, W=H, side=side, e=e, seed=seed, ux=ux, uy=uy, hole=hole)
        pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
        pl = CGPlan(pa)
        pl.set_reference(torch.as_tensor(ref, device="cuda"), "cubic")
        out.append((pl, torch.as_tensor(de, device="cuda")))
    return out


def _run(entry, probs, max_iter=60, tol=1e-4, reg=0.1):
    import torch

    from d2ic_amd import _lib

    P = len(probs)
    vp = ctypes.c_void_p
    arr = lambda xs: (vp * P)(*[vp(x) for x in xs])  # noqa: E731
    us = [torch.zeros((pl.n_nodes, 2), device="cuda") for pl, _ in probs]
    hs = [torch.empty((max_iter, 3), device="cuda") for _ in probs]
    cnts = [torch.zeros(2, dtype=torch.int32, device="cuda") for _ in probs]
    _lib.call(entry, P, arr([pl.handle.value for pl, _ in probs]), arr([g.data_ptr() for _, g in probs]),
              (ctypes.c_int32 * P)(*[g.shape[0] for _, g in probs]), (ctypes.c_int32 * P)(*[g.shape[1] for _, g in probs]),
              arr([u.data_ptr() for u in us]), (ctypes.c_float * P)(*([reg] * P)),
              (ctypes.c_int32 * P)(*([max_iter] * P)), (ctypes.c_float * P)(*([tol] * P)),
              arr([h.data_ptr() for h in hs]), arr([c.data_ptr() for c in cnts]),
              arr([c.data_ptr() + 4 for c in cnts]), _lib.stream())
    torch.cuda.synchronize()
    return [(u.cpu().numpy(), h.cpu().numpy(), c.cpu().numpy()) for u, h, c in zip(us, hs, cnts)]


def test_noalloc_equals_batch_on_two_calls(gpu):
    print("[case] no-alloc solve: 3 problems, then 1 smaller problem", flush=True)
    first = _plans([(128, 96, 16, 1, 0.3, -0.2, 0.0), (160, 128, 8, 2, -0.4, 0.6, 20.0),
                    (160, 128, 32, 4, 0.2, 0.5, 0.0)])
    second = _plans([(96, 64, 16, 3, 0.8, 0.1, 0.0)])
    for probs in (first, second):
        got = _run("d2ic_cg_solve_batch_noalloc", probs)
        want = _run("d2ic_cg_solve_batch", probs)
        for (ug, hg, cg), (uw, hw, cw) in zip(got, want):
            assert cg[0] > 0 and np.array_equal(cg, cw)
            assert np.array_equal(ug, uw)
            assert np.array_equal(hg, hw, equal_nan=True)
        print(f"[case] {len(probs)} problem(s): bitwise equal, iterations {[int(c[0]) for _, _, c in got]}", flush=True)


def test_noalloc_rejects_more_than_eight_problems(gpu):
    from d2ic_amd import _lib

    probs = _plans([(96, 64, 16, 3, 0.8, 0.1, 0.0)]) * 9  # the size check comes first: no device work
    with pytest.raises(ValueError, match="at most|\\[1, 8\\]"):
        _run("d2ic_cg_solve_batch_noalloc", probs)
    assert "8" in _lib.load().d2ic_last_error().decode()
