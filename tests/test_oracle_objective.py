"""Objective, gradient and NLCG of the oracle: closed forms, FD gradients, control-flow semantics."""
from __future__ import annotations

import numpy as np
import pytest

from d2ic_amd import speckle
from d2ic_amd.mask2mesh import mask_to_mesh
from oracle import oracle as O
from tests._cases import structured_case


def _small(interp_noise=0.0, seed=0, ux=0.3, uy=-0.2):
    m, mesh, ref, de = structured_case(H=64, W=64, side=48, e=16, seed=seed, ux=ux, uy=uy, noise=interp_noise)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    return mesh, a, ref, de


def test_reg_energy_zero_for_uniform_and_gradient_is_laplacian():
    mesh, a, ref, _ = _small()
    nn = mesh.nodes_xy.shape[0]
    u = np.tile([0.4, -0.7], (nn, 1))
    J0, g0 = O.objective(u, ref, ref, a, alpha_reg=0.0, interp="linear", dtype=np.float64)
    J1, g1 = O.objective(u, ref, ref, a, alpha_reg=1.0, interp="linear", dtype=np.float64)
    assert abs(J1 - J0) < 1e-12
    np.testing.assert_allclose(g1, g0, atol=1e-12)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(nn, 2))
    _, ga = O.objective(u, ref, ref, a, alpha_reg=0.0, interp="linear", dtype=np.float64)
    _, gb = O.objective(u, ref, ref, a, alpha_reg=2.0, interp="linear", dtype=np.float64)
    idx, deg, w = a["node_neighbor_index"], a["node_neighbor_degree"], a["node_neighbor_weight"]
    L = np.zeros((nn, 2))
    for i in range(nn):
        for k in range(deg[i]):
            j = idx[i, k]
            L[i] += w[i, k] * (u[i] - u[j])
    np.testing.assert_allclose((gb - ga) / 2.0, L, rtol=1e-10, atol=1e-12)


def test_zero_residual_at_true_translation_of_affine_image():
    yy, xx = np.mgrid[0:64, 0:64].astype(np.float64)
    f = 0.9 * xx + 0.4 * yy
    t = (0.25, -0.5)
    g = 0.9 * (xx - t[0]) + 0.4 * (yy - t[1])
    # ROI away from the image border, where clamping breaks affine reproduction
    mesh = mask_to_mesh(speckle.rect_mask(64, 64, 24, 24), 16)
    assert mesh.nodes_xy[mesh.elements].min() >= 16.0
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (64, 64))
    u = np.tile(t, (mesh.nodes_xy.shape[0], 1))
    for interp in ("linear", "cubic"):
        J, grad = O.objective(u, f, g, a, alpha_reg=0.3, interp=interp, dtype=np.float64)
        assert J < 1e-20
        assert np.abs(grad).max() < 1e-10


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_gradient_matches_central_differences(interp):
    mesh, a, ref, de = _small(seed=3)
    rng = np.random.default_rng(1)
    nn = mesh.nodes_xy.shape[0]
    u = rng.normal(scale=0.2, size=(nn, 2))
    J, g = O.objective(u, ref, de, a, alpha_reg=0.1, interp=interp, dtype=np.float64)
    used = np.nonzero(a["node_pixel_degree"] > 0)[0]
    h = 1e-5
    for n in used[::3]:
        for c in range(2):
            up = u.copy(); up[n, c] += h
            um = u.copy(); um[n, c] -= h
            fd = (O.objective(up, ref, de, a, 0.1, interp, np.float64, False) -
                  O.objective(um, ref, de, a, 0.1, interp, np.float64, False)) / (2 * h)
            assert abs(fd - g[n, c]) <= 2e-4 * max(1.0, abs(g[n, c])) + 1e-7, (n, c, fd, g[n, c])


def test_cg_semantics_max_iter_zero_and_nan_padding():
    mesh, a, ref, de = _small()
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    u, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=0, tol=1e-3)
    assert n == 0 and np.all(u == 0) and h.shape == (0, 3)
    u, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=4, tol=1e-12, dtype=np.float64)
    assert n == 4 and np.isfinite(h).all()
    # huge tol: stop at k = 0 before any update, history row 0 = [J0, |g0|, 0], rest NaN
    u, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=5, tol=1e9, dtype=np.float64)
    J0, g0 = O.objective(u0, ref, de, a, 0.1, "cubic", np.float64)
    assert n == 0 and np.all(u == 0)
    np.testing.assert_allclose(h[0], [J0, np.linalg.norm(g0), 0.0], rtol=1e-12)
    assert np.isnan(h[1:]).all()


def test_cg_first_step_is_steepest_descent_with_armijo():
    mesh, a, ref, de = _small()
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2))
    u1, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=1, tol=1e-12, dtype=np.float64)
    J0, g0 = O.objective(u0, ref, de, a, 0.1, "cubic", np.float64)
    alpha = h[0, 2]
    np.testing.assert_allclose(u1, u0 - alpha * g0, atol=1e-14)
    Jc = O.objective(u1, ref, de, a, 0.1, "cubic", np.float64, False)
    assert Jc <= J0 - 1e-4 * alpha * np.sum(g0 * g0)
    assert h[0, 0] == pytest.approx(Jc, rel=1e-12)


def test_cg_recovers_translation_fp32_fp64():
    mesh, a, ref, de = _small(interp_noise=0.5, ux=0.37, uy=-0.52)
    u0 = np.zeros((mesh.nodes_xy.shape[0], 2), np.float32)
    used = a["node_pixel_degree"] > 0
    for dt in (np.float32, np.float64):
        u, h, n = O.cg_solve(u0, ref, de, a, alpha_reg=0.1, max_iter=300, tol=1e-4, dtype=dt)
        m = u[used].mean(0)
        assert abs(m[0] - 0.37) < 0.03 and abs(m[1] + 0.52) < 0.03
        assert np.all(u[~used] == 0)
