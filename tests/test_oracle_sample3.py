"""The reference's real-image known-answer test (tests/test_validation_imageserie_full.py:33-52
with doc/00_Validation_ImageSerie.py's settings) on the oracle: DIC Challenge
Sample3 frames with rigid shifts of 0.1, 0.5 and 1.0 px, Q4-20 structured mesh
over the full image (the reference's roi.tif is all-true), pixel assets without
roi_mask (make_mesh_assets drops it), cubic, max_iters 400, tol 1e-3, reg 0.1,
warm start from the previous frame.  Bar: |mean nodal error| < 0.1 px per frame
and component, as in the reference's test."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import oracle as O

FIX = Path(__file__).resolve().parent / "golden" / "sample3.npz"


def test_oracle_sample3_rigid_shifts():
    from d2ic_amd.mask2mesh import mask_to_mesh

    z = np.load(FIX)
    ref = z["ref"].astype(np.float32)
    frames = z["frames"].astype(np.float32)
    shifts = z["shifts"]
    mesh = mask_to_mesh(np.ones(ref.shape, bool), 20.0, remove_islands=True, min_island_area_px=64)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=None)
    u = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    for k in range(len(frames)):
        u = O.cg_solve(u, ref, frames[k], a, alpha_reg=0.1, max_iter=400, tol=1e-3, interp="cubic")[0]
        err = u.astype(np.float64) - shifts[k]
        print(z["names"][k], "mean error", err.mean(0), "std", err.std(0))
        assert np.all(np.isfinite(err))
        assert np.all(np.abs(err.mean(0)) < 0.1)
