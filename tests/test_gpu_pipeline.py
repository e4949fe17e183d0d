"""Sequence pipeline (BatchMeshBased, batch_mesh_based.py:25-240) with a
global solve chained into the local Gauss-Newton refinement, warm starts from
the previous frame and per-frame NPZ output."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = W = 128


def test_batch_sequence_with_local_refinement(gpu, tmp_path):
    from d2ic_amd import (BatchConfig, BatchMeshBased, DICMeshBased, GlobalCGSolver, LocalGaussNewtonSolver,
                          MeshDICConfig, mask_to_mesh_assets, speckle)

    print("[case] batch pipeline: 3 frames, global CG + 2 local sweeps, NPZ output", flush=True)
    t = gpu
    m = speckle.rect_mask(H, W, 96, 96)
    mesh, assets = mask_to_mesh_assets(m, 16)
    shifts = [(0.2 * k, -0.1 * k) for k in (1, 2, 3)]
    ref, frames = speckle.speckle_sequence_device(
        H, W, 4, [(lambda s: (lambda x, y: (np.full_like(x, s[0]), np.full_like(y, s[1]))))(s) for s in shifts])
    glob = DICMeshBased(mesh, GlobalCGSolver(interpolation="cubic"),
                        MeshDICConfig(max_iters=200, tol=1e-4, reg_strength=0.05, save_history=False))
    loc = DICMeshBased(mesh, LocalGaussNewtonSolver(lam=0.1, max_step=0.2, omega=0.5, interpolation="cubic"),
                       MeshDICConfig(max_iters=2, tol=1e-2, reg_strength=0.05, save_history=True))
    cfg = BatchConfig(save_per_frame=True, per_frame_dir=str(tmp_path))
    batch = BatchMeshBased(ref, assets, glob, cfg, dic_local=loc)
    out = batch.run(frames)
    assert len(out.results) == 3
    nodes_active = None
    for k, (res, s) in enumerate(zip(out.results, shifts)):
        u = res.u_nodal.detach().cpu().numpy()
        assert np.all(np.isfinite(u))
        z = np.load(tmp_path / f"frame_{k:04d}.npz")
        assert set(z.files) >= {"u_nodal", "strain"}
        np.testing.assert_array_equal(z["u_nodal"], u)
        if nodes_active is None:
            nodes_active = np.abs(u).sum(1) > 0
        err = np.abs(u[nodes_active] - np.array(s, np.float32)).max()
        print(f"frame {k}: max |u - shift| = {err:.3f} px")
        assert err < 0.1
    # the chained frame-0 result is the local solver run from the global solution
    glob.set_initial_guess(None)
    r0 = glob.run(frames[0])
    loc.set_initial_guess(r0.u_nodal.clone())
    r1 = loc.run(frames[0])
    np.testing.assert_array_equal(r1.u_nodal.cpu().numpy(), out.results[0].u_nodal.cpu().numpy())
