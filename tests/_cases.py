"""Shared synthetic cases for the tests (small sizes the oracle finishes in seconds)."""
from __future__ import annotations

import numpy as np

from d2ic_amd import speckle
from d2ic_amd.mask2mesh import mask_to_mesh
from d2ic_amd.mesh_assets import Mesh


def structured_case(H=128, W=128, side=96, e=16, seed=0, ux=0.37, uy=-0.52, noise=1.0, hole=0.0):
    if hole > 0:
        m = speckle.plate_with_hole_mask(H, W, side, hole)
    else:
        m = speckle.rect_mask(H, W, side, side)
    mesh = mask_to_mesh(m, e)
    ref, de = speckle.speckle_pair(H, W, seed, speckle.rigid(ux, uy), noise_sigma=noise)
    return m, mesh, ref, de


def distorted_mesh(nx=6, ny=5, h=13.0, x0=7.3, y0=9.1, amp=2.7, seed=3):
    """Non-axis-aligned convex quads (jittered grid) with non-dyadic coordinates."""
    rng = np.random.default_rng(seed)
    xs = x0 + h * np.arange(nx)
    ys = y0 + h * np.arange(ny)
    X, Y = np.meshgrid(xs, ys)
    X = X + rng.uniform(-amp, amp, X.shape)
    Y = Y + rng.uniform(-amp, amp, Y.shape)
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1).astype(np.float32)
    el = []
    for iy in range(ny - 1):
        for ix in range(nx - 1):
            n0 = iy * nx + ix
            el.append([n0, n0 + 1, n0 + nx + 1, n0 + nx])
    return Mesh(nodes_xy=nodes, elements=np.asarray(el, np.int32))
