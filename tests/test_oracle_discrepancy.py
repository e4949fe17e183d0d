"""Pins the oracle's discrepancy-map restatement (discrepancy.py:51-127) with
closed forms: exact samples at pixel centres and on affine images, the
validity edge of the bilinear sampler, the bincount averaging of duplicated
image pixels, and the reference's NaN RMS once a sample is invalid."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def _affine(H, W, a, b, c):
    i, j = np.mgrid[0:H, 0:W].astype(np.float64)
    return (a * j + b * i + c).astype(np.float32)


def test_bilinear_exact_at_centres_and_nan_edge():
    H, W = 6, 9
    im = np.random.default_rng(0).normal(size=(H, W)).astype(np.float32)
    i, j = np.mgrid[0:H, 0:W]
    xy = np.stack([j.ravel() + 0.5, i.ravel() + 0.5], 1).astype(np.float32)
    v = O.bilinear_nan(im, xy).reshape(H, W)
    # the last row and column have no (i0 + 1, j0 + 1) neighbour: outside
    np.testing.assert_array_equal(v[:-1, :-1], im[:-1, :-1])
    assert np.all(np.isnan(v[-1, :])) and np.all(np.isnan(v[:, -1]))
    # just inside / just outside the domain on each side
    pts = np.array([[0.5, 0.5], [0.49, 2.5], [2.5, 0.49], [W - 0.5 - 1e-3, 2.5], [2.5, H - 0.5 - 1e-3]], np.float32)
    s = O.bilinear_nan(im, pts)
    assert np.isfinite(s[0]) and np.isnan(s[1]) and np.isnan(s[2]) and np.isfinite(s[3]) and np.isfinite(s[4])


def test_bilinear_reproduces_affine_images():
    # I[i, j] = a j + b i + c sampled at (x, y) is a (x - .5) + b (y - .5) + c
    # (a != b: a transposed axis fails)
    H, W, a, b, c = 40, 50, 0.75, -1.25, 3.0
    im = _affine(H, W, a, b, c)
    rng = np.random.default_rng(1)
    xy = np.stack([rng.uniform(0.5, W - 0.6, 500), rng.uniform(0.5, H - 0.6, 500)], 1).astype(np.float32)
    v = O.bilinear_nan(im, xy).astype(np.float64)
    want = a * (xy[:, 0].astype(np.float64) - 0.5) + b * (xy[:, 1].astype(np.float64) - 0.5) + c
    np.testing.assert_allclose(v, want, atol=2e-4)


def _toy_assets(coords, flat, H, W):
    n = len(coords)
    return dict(pixel_coords_ref=np.asarray(coords, np.float32).reshape(-1, 2),
                pixel_nodes=np.zeros((n, 4), np.int32),
                pixel_shapeN=np.tile(np.array([1.0, 0.0, 0.0, 0.0], np.float32), (n, 1)),
                roi_mask_flat=np.asarray(flat, np.int64), image_shape=(H, W))


def test_map_duplicates_average_and_invalid_poisons():
    H, W, a, b = 8, 10, 0.5, 2.0
    ref = _affine(H, W, a, b, 1.0)
    de = ref.copy()
    u = np.array([[0.25, -0.5]], np.float32)  # r = a ux + b uy = -0.875 wherever valid
    r0 = a * 0.25 + b * -0.5
    coords = [[3.5, 3.5], [4.5, 3.5], [5.5, 4.5], [W - 0.5, 2.5], [2.5, 5.5]]
    # pixels 0 and 1 share image pixel 7; pixel 3 (last column: invalid) shares 12 with pixel 4
    flat = [7, 7, 20, 12, 12]
    m, rms = O.discrepancy_map(ref, de, u, _toy_assets(coords, flat, H, W))
    assert abs(m.ravel()[7] - r0) < 1e-6
    assert abs(m.ravel()[20] - r0) < 1e-6
    assert np.isnan(m.ravel()[12])
    others = np.ones(H * W, bool)
    others[[7, 20, 12]] = False
    assert np.all(np.isnan(m.ravel()[others]))
    assert np.isnan(rms)  # the invalid pixel makes the reference's RMS NaN
    # without the invalid pixel the RMS is finite, and the duplicates average
    de2 = de.copy()
    de2[2, 5] += 1.0  # pixel 1 samples at (4.75, 3.0): taps rows 2-3, cols 4-5; (2, 5) has weight 1/8
    m2, rms2 = O.discrepancy_map(ref, de2, u, _toy_assets(coords[:3], flat[:3], H, W))
    r1 = O.bilinear_nan(de2, np.array([[4.75, 3.0]], np.float32))[0] - O.bilinear_nan(ref, np.array([[4.5, 3.5]],
                                                                                                    np.float32))[0]
    assert abs(r1 - r0) > 0.1
    assert abs(m2.ravel()[7] - np.float32(0.5) * (np.float32(r0) + r1)) < 1e-6
    assert abs(rms2 - np.sqrt((2 * r0 ** 2 + r1 ** 2) / 3)) < 1e-6


def test_map_on_structured_assets_and_empty():
    from d2ic_amd import speckle
    from d2ic_amd.mask2mesh import mask_to_mesh

    H = W = 64
    m = speckle.rect_mask(H, W, 40, 40)
    mesh = mask_to_mesh(m, 8)
    a = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=m)
    ref = _affine(H, W, 0.3, -0.7, 2.0)
    u = np.tile(np.array([[0.4, 0.1]], np.float32), (mesh.nodes_xy.shape[0], 1))
    mp, rms = O.discrepancy_map(ref, ref, u, a)
    want = 0.3 * 0.4 - 0.7 * 0.1
    inside = np.zeros(H * W, bool)
    inside[np.asarray(a["roi_mask_flat"])] = True
    inside = inside.reshape(H, W)
    assert inside.sum() == len(a["roi_mask_flat"]) > 1000  # distinct image pixels
    assert np.all(np.isnan(mp[~inside]))
    np.testing.assert_allclose(mp[inside], want, atol=2e-5)
    assert abs(rms - abs(want)) < 2e-5
    empty = dict(pixel_coords_ref=np.zeros((0, 2), np.float32), pixel_nodes=np.zeros((0, 4), np.int32),
                 pixel_shapeN=np.zeros((0, 4), np.float32), roi_mask_flat=np.zeros(0, np.int64), image_shape=(H, W))
    me, rmse = O.discrepancy_map(ref, ref, u, empty)
    assert np.all(np.isnan(me)) and rmse == 0.0
