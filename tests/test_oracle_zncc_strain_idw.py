"""ZNCC / strain / IDW oracle vs the reference's fixtures, numpy brute force and closed forms."""
from __future__ import annotations

import numpy as np

from d2ic_amd import Mesh, build_node_neighbor_tables, mask_to_mesh
from oracle import oracle as O


def test_zncc_reference_fixture():
    """tests/test_solver_translation_zncc.py:8-33 of the reference."""
    rng = np.random.default_rng(0)
    ref = rng.normal(size=(64, 64)).astype(np.float32)
    de = np.zeros_like(ref)
    de[2:, 3:] = ref[:-2, :-3]
    centers = np.array([[32.0, 32.0], [40.0, 24.0], [24.0, 40.0]], np.float32)
    u, s = O.zncc(centers, ref, de, 9, 6)
    assert np.isfinite(s).all()
    assert np.allclose(u, [3, 2], atol=0.5)


def _zncc_numpy(c, ref, de, win, s):
    x, y = int(np.rint(c[0])), int(np.rint(c[1]))
    r = win // 2
    t = ref[y - r:y + r + 1, x - r:x + r + 1].astype(np.float64)
    t = (t - t.mean()) / (t.std() + 1e-12)
    best, bi = -np.inf, 0
    n = 2 * s + 1
    for ky in range(n):
        for kx in range(n):
            w = de[y - s - r + ky:y - s - r + ky + win, x - s - r + kx:x - s - r + kx + win].astype(np.float64)
            w = (w - w.mean()) / (w.std() + 1e-12)
            sc = (w * t).mean()
            if sc > best:
                best, bi = sc, ky * n + kx
    return np.array([bi % n - s, bi // n - s]), best


def test_zncc_vs_numpy_bruteforce():
    rng = np.random.default_rng(4)
    ref = rng.normal(size=(80, 80)).astype(np.float32)
    de = np.roll(ref, (1, -2), axis=(0, 1)) + 0.3 * rng.normal(size=ref.shape).astype(np.float32)
    centers = np.array([[30.0, 30.0], [45.4, 38.6], [50.5, 41.5], [2.0, 2.0]], np.float32)
    u, s = O.zncc(centers, ref, de, 11, 5, dtype=np.float64)
    for k in range(3):
        ub, sb = _zncc_numpy(centers[k], ref, de, 11, 5)
        assert np.array_equal(u[k], ub)
        assert abs(s[k] - sb) < 1e-10
    assert np.isnan(s[3]) and np.all(u[3] == 0)


def test_strain_reference_closed_form():
    """tests/test_strain_smoke.py:10-40 of the reference."""
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]], np.float32)
    mesh = Mesh(nodes_xy=nodes, elements=np.array([[0, 1, 3, 2]], np.int32))
    idx, deg = build_node_neighbor_tables(mesh)
    eps = 0.01
    disp = np.stack([nodes[:, 0] * eps, np.zeros(4, np.float32)], 1)
    _, _, v = O.strain(disp, nodes, idx, deg)
    np.testing.assert_allclose(v[:, 0], eps + 0.5 * eps ** 2, atol=5e-4)
    np.testing.assert_allclose(v[:, 1:], 0.0, atol=5e-4)


def test_strain_exact_for_affine_fields():
    mesh = mask_to_mesh(np.ones((80, 96), bool), 16)
    X = mesh.nodes_xy.astype(np.float64)
    G = np.array([[0.02, -0.01], [0.005, -0.03]])
    u = X @ G.T + [1.5, -0.25]
    idx, deg = build_node_neighbor_tables(mesh)
    F, E, v = O.strain(u, X, idx, deg, eps=0.0, dtype=np.float64)
    Fx = np.eye(2) + G
    Ex = 0.5 * (Fx.T @ Fx - np.eye(2))
    np.testing.assert_allclose(F, np.broadcast_to(Fx, F.shape), atol=1e-12)
    np.testing.assert_allclose(E, np.broadcast_to(Ex, E.shape), atol=1e-12)
    np.testing.assert_allclose(v[:, 2], 0.5 * (Ex[0, 1] + Ex[1, 0]), atol=1e-12)


def test_idw_closed_form_and_exact_hit():
    nodes = np.array([[0.0, 0.0], [2.0, 0.0], [1.0, 1.0]], np.float64)
    centers = np.array([[1.0, 0.0], [1.0, 1.0]], np.float64)
    disp = np.array([[1.0, 2.0], [3.0, 4.0]], np.float64)
    out = O.idw(nodes, centers, disp, dtype=np.float64)
    # node 0: distances 1 and sqrt(2) -> weights 1 and 1/2
    np.testing.assert_allclose(out[0], (np.array([1, 2]) * 1 + np.array([3, 4]) * 0.5) / 1.5, rtol=1e-12)
    assert np.array_equal(out[2], disp[1])
    nan_disp = disp.copy()
    nan_disp[1] = np.nan
    out2 = O.idw(nodes, centers, nan_disp, dtype=np.float64)
    assert np.array_equal(out2[0], disp[0])
