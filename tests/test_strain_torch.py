"""The differentiable strain restatement used by the adjoint losses
(``green_lagrange_torch``, reference ``strain.py:11-114``) on the CPU: values
against the oracle's strain (FP64) and gradients (including batched u and the
Gaussian gauge weights) against central finite differences, which pins the
custom fixed-order backward of its neighbour gather."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from d2ic_amd import _lib, speckle
from d2ic_amd.mask2mesh import mask_to_mesh
from d2ic_amd.mesh_assets import build_node_neighbor_tables
from oracle import oracle as O


@pytest.fixture(autouse=True)
def cpu_device(monkeypatch):
    monkeypatch.setattr(_lib, "device", lambda: torch.device("cpu"))


@pytest.mark.parametrize("gauge", [0.0, 20.0])
def test_values_and_gradients(gauge):
    from d2ic_amd.adjoint import green_lagrange_torch

    mesh = mask_to_mesh(speckle.plate_with_hole_mask(128, 128, 110, 20.0), 16)
    idx, deg = build_node_neighbor_tables(mesh)
    X = np.asarray(mesh.nodes_xy, np.float64)
    rng = np.random.default_rng(3)
    u0 = rng.normal(scale=0.3, size=(2, X.shape[0], 2))
    E = green_lagrange_torch(torch.tensor(u0), X, idx, deg, gauge_length=gauge).numpy()
    for b in range(2):
        _, Eo, _ = O.strain(u0[b], X, np.asarray(idx), np.asarray(deg), gauge_length=gauge, dtype=np.float64)
        np.testing.assert_allclose(E[b], Eo, atol=1e-12)
    wts = rng.normal(size=E.shape)
    u = torch.tensor(u0, requires_grad=True)
    loss = (green_lagrange_torch(u, X, idx, deg, gauge_length=gauge) * torch.tensor(wts)).sum()
    (g,) = torch.autograd.grad(loss, u)
    g = g.numpy()

    def f(v):
        return float((green_lagrange_torch(torch.tensor(v), X, idx, deg, gauge_length=gauge).numpy() * wts).sum())

    h = 1e-6
    for b, n, c in [(0, 0, 0), (1, 5, 1), (0, 17, 0), (1, X.shape[0] - 1, 1), (0, 40, 1)]:
        up, um = u0.copy(), u0.copy()
        up[b, n, c] += h
        um[b, n, c] -= h
        fd = (f(up) - f(um)) / (2 * h)
        assert abs(fd - g[b, n, c]) <= 1e-6 * max(1.0, abs(fd)), (b, n, c, fd, g[b, n, c])
