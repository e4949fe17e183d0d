"""Persistent grid size (d2ic_cg_plan_set_grid) and the grid autotuner.

Every reduction of the persistent NLCG kernel is an order-independent
fixed-point sum, so the solve is bitwise the same whatever the grid: u, the
[J, |grad|, alpha] history and the iteration count are compared exactly for
the work-based default grid and forced grids of 8, 64 and 256 blocks.  The
autotuner must return one of its candidates, keep it on the plan, and leave
results unchanged."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_solve_is_bitwise_identical_across_grids(gpu, tmp_path):
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets, speckle

    H = W = 512
    print("[case] grid sizes 0/8/64/256 on 512^2 Q4-16, then autotune", flush=True)
    m = speckle.plate_with_hole_mask(H, W, 440, 70.0)
    mesh, assets = mask_to_mesh_assets(m, 16)
    ref, de = speckle.speckle_pair(H, W, 5, speckle.rigid(0.45, -0.3), noise_sigma=0.5)
    solver = GlobalCGSolver(interpolation="cubic")
    dic = DICMeshBased(mesh, solver, MeshDICConfig(max_iters=60, tol=1e-5, reg_strength=0.1, save_history=True))
    dic.prepare(ref, assets)
    out = {}
    for b in (0, 8, 64, 256):
        solver.plan.set_grid(b)
        dic.set_initial_guess(None)
        res = dic.run(de)
        out[b] = (res.u_nodal.cpu().numpy(), res.history.cpu().numpy(), res.diagnostics.info["n_iters"],
                  solver.last_pass_count)
        print(f"grid {b}: iters {out[b][2]} passes {out[b][3]} kernel {solver.plan.kernel_ms():.3f} ms")
    u0, h0, n0, p0 = out[0]
    assert 0 < n0 <= 60
    for b in (8, 64, 256):
        u, h, n, p = out[b]
        assert n == n0 and p == p0
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(h, h0)

    cache = tmp_path / "grid_autotune.json"
    tune = solver.autotune_grid(dic.state, de, candidates=(16, 64, 128), max_iter=6, repeats=2, cache=str(cache))
    print(f"autotune {tune}")
    assert tune["blocks"] in (0, 16, 64, 128) and set(tune["ms"]) == {0, 16, 64, 128}
    assert all(v > 0.0 for v in tune["ms"].values()) and tune["default_ms"] == tune["ms"][0]
    assert not tune["cached"] and cache.exists()
    assert solver.plan.grid_blocks == tune["blocks"]
    # the persisted choice is applied without timing
    solver.plan.set_grid(0)
    again = solver.autotune_grid(dic.state, de, candidates=(16, 64, 128), max_iter=6, repeats=2, cache=str(cache))
    assert again["cached"] and again["blocks"] == tune["blocks"] and solver.plan.grid_blocks == tune["blocks"]
    dic.set_initial_guess(None)
    res = dic.run(de)
    np.testing.assert_array_equal(res.u_nodal.cpu().numpy(), u0)
    np.testing.assert_array_equal(res.history.cpu().numpy(), h0)


def test_tilings_solve_the_same_problem_and_autotune(gpu, tmp_path):
    """Plans with other tilings (tiles closed at 64/128/192 pixels, 8- or
    12-pixel strips) solve the same problem: only the grouping of the tile
    partials changes, so u agrees with the default tiling to rounding (bar
    1e-5 px after 30 iterations, tol 0).  The tiling autotuner keeps one of
    its candidates as the solver's plan and re-applies a cached choice."""
    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, mask_to_mesh_assets, speckle

    H = W = 512
    for e in (20, 32):
        print(f"[case] tilings on 512^2 plate Q4-{e}", flush=True)
        m = speckle.plate_with_hole_mask(H, W, 440, 70.0)
        mesh, assets = mask_to_mesh_assets(m, e)
        ref, de = speckle.speckle_pair(H, W, 9, speckle.rigid(0.31, -0.42), noise_sigma=0.5)
        solver = GlobalCGSolver(interpolation="cubic")
        dic = DICMeshBased(mesh, solver, MeshDICConfig(max_iters=30, tol=0.0, reg_strength=0.1))
        dic.prepare(ref, assets)
        res = dic.run(de)
        u_def = res.u_nodal.cpu().numpy()
        tiles_def = solver.plan.info()["n_tiles"]
        assert solver.plan.tiling == (256, 16)
        for tp, sw in ((128, 16), (64, 16), (256, 8), (192, 12)):
            solver.set_tiling(dic.state, tp, sw)
            info = solver.plan.info()
            dic.set_initial_guess(None)
            u = dic.run(de).u_nodal.cpu().numpy()
            d = float(np.abs(u - u_def).max())
            print(f"tiling {tp}/{sw}: tiles {info['n_tiles']} (default {tiles_def}) max|du| {d:.2e}")
            assert solver.plan.tiling == (tp, sw) and info["n_pixels"] == solver.plan.n_pixels
            assert d <= 1e-5
            if tp == 64:
                assert info["n_tiles"] > tiles_def
        solver.set_tiling(dic.state)
        cache = tmp_path / f"tiling_{e}.json"
        tune = solver.autotune_tiling(dic.state, de, candidates=[(128, 16), (256, 8)], max_iter=6, repeats=2,
                                      cache=str(cache))
        print(f"autotune {tune}")
        assert tune["tiling"] in ((256, 16), (128, 16), (256, 8)) and not tune["cached"]
        assert set(tune["ms"]) == {"256/16", "128/16", "256/8"} and tune["default_ms"] == tune["ms"]["256/16"]
        assert solver.plan.tiling == tune["tiling"] and cache.exists()
        solver.set_tiling(dic.state)
        again = solver.autotune_tiling(dic.state, de, candidates=[(128, 16), (256, 8)], max_iter=6, repeats=2,
                                       cache=str(cache))
        assert again["cached"] and again["tiling"] == tune["tiling"] and solver.plan.tiling == tune["tiling"]
        dic.set_initial_guess(None)
        u = dic.run(de).u_nodal.cpu().numpy()
        assert float(np.abs(u - u_def).max()) <= 1e-5
