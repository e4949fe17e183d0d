"""TranslationZNCCSolver.solve and DICInitMotion.run on the GPU against the
oracle: score_min rejection (u = 0, score NaN; ``solver_translation_zncc.py:74-77``),
max_centers truncation (``:59-61``), out-of-bounds centres (u = 0, score
NaN), flat templates (every score 0: the first window, u = (-s, -s), wins),
and DICInitMotion's diagnostics and IDW projection
(``dic_init_motion.py:41-91``; rejected centres hold u = 0 and still vote)."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

H = W = 160
WIN, SEARCH = 21, 4


def _images():
    from d2ic_amd import speckle

    ref, de = speckle.speckle_pair(H, W, 31, speckle.rigid(2.0, -1.0), noise_sigma=1.0)
    de = de.copy()
    rng = np.random.default_rng(2)
    de[:, 110:] = rng.normal(128.0, 40.0, size=(H, W - 110)).astype(np.float32)  # unrelated texture: low scores
    ref = ref.copy()
    ref[8:40, 8:40] = 77.0  # a flat region of the reference: flat templates
    return ref, de


def _centres():
    from d2ic_amd.mesh_assets import compute_element_centers
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd import speckle

    mesh = mask_to_mesh(speckle.rect_mask(H, W, 128, 128), 16)
    c = np.asarray(compute_element_centers(mesh), np.float32)
    extra = np.array([[24.0, 24.0],  # flat template, patch inside
                      [5.0, 80.0],   # template leaves the image
                      [80.0, 154.5],  # search patch leaves the image
                      [80.5, 80.5]], np.float32)  # half-way centre: rint rounds half to even
    return mesh, np.concatenate([c, extra])


def _expected(c, ref, de, score_min):
    u, s = O.zncc(c, ref, de, WIN, SEARCH)
    if score_min is not None:
        bad = ~(s >= score_min)
        u[bad] = 0.0
        s[bad] = np.nan
    return u, s


@pytest.mark.parametrize("score_min,max_centers", [(None, None), (0.5, None), (0.5, 40), (0.2, 200)])
def test_zncc_solver_solve(gpu, score_min, max_centers):
    from d2ic_amd import InitMotionConfig, TranslationZNCCSolver

    t = gpu
    print(f"[case] ZNCC solve score_min={score_min} max_centers={max_centers}", flush=True)
    ref, de = _images()
    _, c = _centres()
    cfg = InitMotionConfig(win=WIN, search=SEARCH, score_min=score_min, max_centers=max_centers)
    solver = TranslationZNCCSolver(cfg)
    state = SimpleNamespace(assets=SimpleNamespace(element_centers_xy=c), ref_image=t.as_tensor(ref, device="cuda"))
    with pytest.raises(RuntimeError):
        solver.solve(state, de)  # before compile()
    solver.compile(None)
    res = solver.solve(state, t.as_tensor(de, device="cuda"))
    n = len(c) if max_centers is None else min(len(c), max_centers)
    u_exp, s_exp = _expected(c[:n], ref, de, score_min)
    u, s = res.u_centers.cpu().numpy(), res.scores.cpu().numpy()
    assert u.shape == (n, 2) and s.shape == (n,)
    np.testing.assert_array_equal(u, u_exp)
    np.testing.assert_array_equal(np.isnan(s), np.isnan(s_exp))
    np.testing.assert_allclose(s, s_exp, rtol=1e-5, atol=1e-6, equal_nan=True)
    if max_centers is None:
        ne = len(c) - 4
        flat, oob_t, oob_p = ne, ne + 1, ne + 2
        assert np.isnan(s[oob_t]) and np.isnan(s[oob_p]) and (u[[oob_t, oob_p]] == 0).all()
        if score_min is None:
            assert s[flat] == 0.0 and tuple(u[flat]) == (-SEARCH, -SEARCH)
            good = s[:ne] > 0.9
            assert good.sum() > 20 and (u[:ne][good] == [2.0, -1.0]).all()
        else:
            assert np.isnan(s[flat]) and tuple(u[flat]) == (0.0, 0.0)
            assert np.isnan(s[:ne]).sum() > 5  # the noise half is rejected


@pytest.mark.parametrize("max_centers", [None, 60])
def test_init_motion_run(gpu, max_centers):
    """Diagnostics keys as the reference sets them; u_nodal = IDW of the
    centre shifts, rejected centres voting with u = 0.  With max_centers below
    the number of centres the reference passes every centre with the shorter
    shift array to its IDW, whose boolean mask then has the wrong length and
    raises; this build projects the first max_centers centres instead."""
    from d2ic_amd import DICInitMotion, InitMotionConfig, MeshAssets, TranslationZNCCSolver
    from d2ic_amd.mesh_assets import compute_element_centers

    t = gpu
    print(f"[case] DICInitMotion.run max_centers={max_centers}", flush=True)
    ref, de = _images()
    mesh, _ = _centres()
    cen = np.asarray(compute_element_centers(mesh), np.float32)
    assets = MeshAssets(mesh=mesh, element_centers_xy=cen)
    cfg = InitMotionConfig(win=WIN, search=SEARCH, score_min=0.5, max_centers=max_centers)
    init = DICInitMotion(cfg, TranslationZNCCSolver(cfg))
    with pytest.raises(RuntimeError):
        init.run(de)
    init.prepare(ref, assets)
    res = init.run(t.as_tensor(de, device="cuda"))
    n = len(cen) if max_centers is None else max_centers
    u_c, s_c = _expected(cen[:n], ref, de, 0.5)
    assert res.diagnostics.info == {"stage": "init_motion", "matches": int((~np.isnan(s_c)).sum()),
                                    "note": "IDW projection"}
    u_exp = O.idw(mesh.nodes_xy, cen[:n], u_c)
    np.testing.assert_allclose(res.u_nodal.cpu().numpy(), u_exp, rtol=1e-5, atol=1e-6)
    assert tuple(res.strain.shape) == (mesh.nodes_xy.shape[0], 2) and float(res.strain.abs().max()) == 0.0
    assert res.history is None
