"""Pins the oracle's local Gauss-Newton restatement (solver_local_gn.py:186-414)
with closed forms on hand-built node tables and a synthetic translation."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def test_image_gradient_ramp():
    H, W = 7, 9
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    gx, gy = O.image_gradient(3.0 * x + 5.0 * y)
    assert np.all(gx[:, 1:-1] == 3.0) and np.all(gx[:, [0, -1]] == 0.0)
    assert np.all(gy[1:-1, :] == 5.0) and np.all(gy[[0, -1], :] == 0.0)


def _tables(npix_per_node, nn, nb=None, w=None, nrw=None):
    """Node->pixel table where node n owns pixels n*k .. n*k+k-1 with N = 1."""
    k = npix_per_node
    a = dict(
        node_pixel_index=np.arange(nn * k, dtype=np.int32).reshape(nn, k),
        node_N_weight=np.ones((nn, k), np.float32),
        node_pixel_degree=np.full(nn, k, np.int32),
        node_neighbor_index=np.zeros((nn, 0), np.int32) if nb is None else np.asarray(nb, np.int32),
        node_neighbor_weight=np.zeros((nn, 0), np.float32) if w is None else np.asarray(w, np.float32),
        node_reg_weight=np.ones(nn, np.float32) if nrw is None else np.asarray(nrw, np.float32),
    )
    a["node_neighbor_degree"] = np.full(nn, a["node_neighbor_index"].shape[1], np.int32)
    return a


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_step_recovers_linear_displacement(dtype):
    # r_p = -(gx_p dx_n + gy_p dy_n) on node n's pixels (N = 1): one undamped
    # step (lam = 0, omega = 1, no clip) moves u by exactly d (up to the 1e-8 I)
    rng = np.random.default_rng(0)
    nn, k = 5, 12
    a = _tables(k, nn)
    gx = rng.normal(size=nn * k).astype(dtype)
    gy = rng.normal(size=nn * k).astype(dtype)
    d = rng.normal(scale=0.3, size=(nn, 2)).astype(dtype)
    owner = np.repeat(np.arange(nn), k)
    r = -(gx * d[owner, 0] + gy * d[owner, 1])
    u0 = rng.normal(size=(nn, 2)).astype(dtype)
    u1 = O.local_gn_step(u0, r, gx, gy, a, lam=0.0, max_step=1e9, alpha_reg=0.0, omega=1.0, dtype=dtype)
    tol = 1e-5 if dtype == np.float32 else 1e-7
    assert np.abs(u1 - u0 - d).max() < tol


def test_lam_omega_and_clip():
    # isotropic H = c I (gradients (1,0) and (0,1) on alternate pixels):
    # delta = d / (1 + 2 lam); omega scales it; max_step clips its norm
    nn, k = 1, 2
    a = _tables(k, nn)
    gx = np.array([1.0, 0.0], np.float32)
    gy = np.array([0.0, 1.0], np.float32)
    d = np.array([0.04, -0.03], np.float32)
    r = -(gx * d[0] + gy * d[1])
    u0 = np.zeros((1, 2), np.float32)
    for lam in (0.0, 0.1, 0.5):
        u1 = O.local_gn_step(u0, r, gx, gy, a, lam=lam, max_step=1e9, alpha_reg=0.0, omega=0.5)
        np.testing.assert_allclose(u1[0], 0.5 * d / (1 + 2 * lam), rtol=1e-5, atol=1e-9)
    big = 10.0 * d
    rb = -(gx * big[0] + gy * big[1])
    u1 = O.local_gn_step(u0, rb, gx, gy, a, lam=0.0, max_step=0.2, alpha_reg=0.0, omega=1.0)
    assert abs(np.linalg.norm(u1[0]) - 0.2) < 1e-6
    np.testing.assert_allclose(u1[0] / np.linalg.norm(u1[0]), big / np.linalg.norm(big), rtol=1e-6)


def test_spring_only_node_moves_to_weighted_mean():
    # node 0 has no pixels and springs to nodes 1, 2 (weights 1, 3): with omega
    # = 1 it moves to (u1 + 3 u2) / 4; node 3 has neither and stays
    a = dict(
        node_pixel_index=np.zeros((4, 1), np.int32), node_N_weight=np.zeros((4, 1), np.float32),
        node_pixel_degree=np.zeros(4, np.int32),
        node_neighbor_index=np.array([[1, 2], [0, 0], [0, 0], [0, 0]], np.int32),
        node_neighbor_weight=np.array([[1.0, 3.0], [0, 0], [0, 0], [0, 0]], np.float32),
        node_neighbor_degree=np.array([2, 0, 0, 0], np.int32),
        node_reg_weight=np.full(4, 2.0, np.float32),
    )
    u0 = np.array([[0.0, 0.0], [0.1, 0.0], [0.05, 0.02], [0.3, 0.3]], np.float32)
    r = np.zeros(1, np.float32)
    u1 = O.local_gn_step(u0, r, r, r, a, lam=0.1, max_step=1e9, alpha_reg=0.5, omega=1.0)
    np.testing.assert_allclose(u1[0], (u0[1] + 3 * u0[2]) / 4, rtol=1e-5)
    np.testing.assert_array_equal(u1[1:], u0[1:])
    u2 = O.local_gn_step(u0, r, r, r, a, lam=0.1, max_step=1e9, alpha_reg=0.0, omega=1.0)
    np.testing.assert_array_equal(u2, u0)


def test_sweeps_recover_translation():
    from d2ic_amd import speckle
    from d2ic_amd.mask2mesh import mask_to_mesh

    H = W = 80
    m = speckle.rect_mask(H, W, 48, 48)
    mesh = mask_to_mesh(m, 16)
    shift = (0.3, -0.2)
    ref, de = speckle.speckle_pair(H, W, 7, lambda x, y: (np.full_like(x, shift[0]), np.full_like(y, shift[1])),
                                   noise_sigma=0.0)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=m)
    nn = a["node_pixel_degree"].shape[0]
    u, hist = O.local_gn_sweeps(np.zeros((nn, 2)), np.asarray(ref), np.asarray(de), a, lam=1e-3, max_step=0.2,
                                alpha_reg=0.0, omega=0.5, n_sweeps=12, interp="cubic")
    act = a["node_pixel_degree"] > 0
    err = np.abs(u[act] - np.array(shift, np.float32)).max()
    assert err < 0.03, err
    assert hist[-1, 0] < 0.5 * hist[0, 0]
    assert np.all(np.isfinite(hist))
