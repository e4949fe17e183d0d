"""Committed regression fixtures (tests/golden/, written by make_golden.py from the oracle).

CPU tests: the oracle still reproduces the fixtures.  GPU tests: the HIP path on
the fixture inputs against the fixture outputs (integer results bit-exact, CG
histories within the same bars as the live-oracle tests)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLD = Path(__file__).resolve().parent / "golden"


def _load(name):
    return dict(np.load(GOLD / name, allow_pickle=False))


def _cg_assets(d):
    return O.pixel_assets(d["nodes"], d["elements"], d["ref"].shape, roi_mask=d["roi_mask"].astype(bool))


def test_oracle_reproduces_assets_fixture():
    d = _load("assets_distorted.npz")
    a = O.pixel_assets(d["nodes"], d["elements"], tuple(d["image_shape"]))
    for k in ("pixel_nodes", "pixel_elts", "node_neighbor_index", "node_neighbor_degree", "node_pixel_index",
              "node_pixel_degree", "roi_mask_flat", "pixel_coords_ref", "pixel_shapeN", "node_N_weight",
              "node_neighbor_weight"):
        assert np.array_equal(np.asarray(a[k]), d[k]), k


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_oracle_reproduces_cg_fixture(interp):
    d = _load("cg_rect64.npz")
    a = _cg_assets(d)
    u0 = np.zeros_like(d["u_cubic_f32"])
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        u, h, n = O.cg_solve(u0, d["ref"], d["de"], a, alpha_reg=float(d["reg"]), max_iter=int(d["max_iter"]),
                             tol=float(d["tol"]), interp=interp, dtype=dt)
        assert n == int(d[f"n_{interp}_{tag}"])
        np.testing.assert_array_equal(h, d[f"hist_{interp}_{tag}"])
        np.testing.assert_array_equal(u, d[f"u_{interp}_{tag}"])


def test_oracle_reproduces_zncc_fixture():
    d = _load("zncc_speckle96.npz")
    u, s = O.zncc(d["centers"], d["ref"], d["de"], int(d["win"]), int(d["search"]))
    assert np.array_equal(u, d["u"])
    np.testing.assert_array_equal(s, d["scores"])


@pytest.mark.gpu
def test_gpu_assets_match_fixture(gpu):
    from d2ic_amd.mesh_assets import Mesh
    from d2ic_amd.pixel_assets import build_pixel_assets

    print("[case] golden assets", flush=True)
    d = _load("assets_distorted.npz")
    mesh = Mesh(nodes_xy=d["nodes"], elements=d["elements"])
    pa = build_pixel_assets(mesh, np.zeros(tuple(d["image_shape"]), np.float32), 1.0)
    for k in ("pixel_nodes", "node_neighbor_index", "node_neighbor_degree", "node_pixel_index", "node_pixel_degree",
              "roi_mask_flat"):
        a = getattr(pa, k).cpu().numpy()
        assert np.array_equal(a, d[k].astype(a.dtype)), k
    assert np.array_equal(pa.pixel_elements.cpu().numpy(), d["pixel_elts"])
    # non-dyadic geometry: Newton results within 4e-7 (FMA contraction differs from the oracle's)
    np.testing.assert_allclose(pa.pixel_shapeN.cpu().numpy(), d["pixel_shapeN"], rtol=0, atol=4e-7)
    np.testing.assert_array_equal(pa.pixel_coords_ref.cpu().numpy(), d["pixel_coords_ref"])


@pytest.mark.gpu
@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_gpu_cg_matches_fixture(gpu, interp):
    """20 iterations: same alpha sequence as the FP64 fixture, J within rtol 2e-5,
    |grad| within rtol 2e-4, u within 2e-5 px."""
    from dataclasses import replace

    from d2ic_amd import DICMeshBased, GlobalCGSolver, MeshDICConfig, make_mesh_assets
    from d2ic_amd.mesh_assets import Mesh

    print(f"[case] golden cg {interp}", flush=True)
    d = _load("cg_rect64.npz")
    mesh = Mesh(nodes_xy=d["nodes"], elements=d["elements"])
    cfg = MeshDICConfig(max_iters=int(d["max_iter"]), tol=float(d["tol"]), reg_strength=float(d["reg"]),
                        save_history=True)
    assets = replace(make_mesh_assets(mesh), roi_mask=d["roi_mask"].astype(bool))
    dic = DICMeshBased(mesh, GlobalCGSolver(interpolation=interp), cfg)
    dic.prepare(d["ref"], assets)
    res = dic.run(d["de"])
    h = res.history.cpu().numpy()
    h64 = d[f"hist_{interp}_f64"]
    assert res.diagnostics.info["n_iters"] == int(d[f"n_{interp}_f64"])
    np.testing.assert_array_equal(h[:, 2], h64[:, 2].astype(np.float32))
    np.testing.assert_allclose(h[:, 0], h64[:, 0], rtol=2e-5)
    np.testing.assert_allclose(h[:, 1], h64[:, 1], rtol=2e-4)
    np.testing.assert_allclose(res.u_nodal.cpu().numpy(), d[f"u_{interp}_f64"], atol=2e-5)


@pytest.mark.gpu
def test_gpu_zncc_matches_fixture(gpu):
    from d2ic_amd.solver_translation_zncc import match_centers

    print("[case] golden zncc", flush=True)
    d = _load("zncc_speckle96.npz")
    u, s = match_centers(d["centers"], d["ref"], d["de"], win=int(d["win"]), search=int(d["search"]))
    assert np.array_equal(u.cpu().numpy(), d["u"])
    np.testing.assert_allclose(s.cpu().numpy(), d["scores"], rtol=1e-5)
