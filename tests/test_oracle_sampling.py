"""Samplers: linear pinned by the reference's own formula; cubic (Keys a=-1/2) by its invariants."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def _affine(H, W, a=0.7, b=-1.3, c=5.0):
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    return (a * xx + b * yy + c)


@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_interpolates_pixel_centres(interp, dtype):
    rng = np.random.default_rng(0)
    shape = (20, 20) if interp == "linear" else (20, 24)
    img = rng.uniform(0, 255, shape).astype(dtype)
    ii, jj = np.mgrid[0:shape[0], 0:shape[1]]
    xy = np.stack([jj.ravel() + 0.5, ii.ravel() + 0.5], 1)
    v = O.sample(img, xy, interp=interp, dtype=dtype)
    assert np.array_equal(v, img.ravel())


def test_linear_non_square_clip_quirk():
    """_bilinear_sample takes its clip bounds from the transposed image
    (solver_global_cg.py:491-496): the column index is clipped by H-1, so for
    H < W columns >= H read column H-1."""
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 255, (10, 14))
    xy = np.array([[12.5, 3.5], [9.5, 3.5]])
    v = O.sample(img, xy, interp="linear", dtype=np.float64)
    assert v[0] == img[3, 9] and v[1] == img[3, 9]


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_reproduces_affine_interior(interp):
    img = _affine(30, 30)
    rng = np.random.default_rng(1)
    xy = rng.uniform(3, 26, (500, 2))
    v, gx, gy = O.sample(img, xy, interp=interp, dtype=np.float64, with_grad=True)
    exact = 0.7 * (xy[:, 0] - 0.5) - 1.3 * (xy[:, 1] - 0.5) + 5.0
    np.testing.assert_allclose(v, exact, atol=1e-10)
    np.testing.assert_allclose(gx, 0.7, atol=1e-10)
    np.testing.assert_allclose(gy, -1.3, atol=1e-10)


def test_cubic_reproduces_quadratic_and_is_not_linear():
    yy, xx = np.mgrid[0:30, 0:30].astype(np.float64)
    img = 0.05 * xx ** 2
    xy = np.array([[10.3 + 0.5, 7.5], [12.75 + 0.5, 9.5]])
    v = O.sample(img, xy, interp="cubic", dtype=np.float64)
    np.testing.assert_allclose(v, 0.05 * (xy[:, 0] - 0.5) ** 2, atol=1e-12)
    vl = O.sample(img, xy, interp="linear", dtype=np.float64)
    assert np.all(np.abs(vl - v) > 1e-3)


@pytest.mark.parametrize("interp", ["linear", "cubic"])
def test_derivative_matches_central_differences(interp):
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 255, (24, 24))
    xy = rng.uniform(4, 20, (300, 2))
    # keep away from the kinks of the bilinear interpolant
    frac = (xy - 0.5) % 1.0
    xy = xy[(np.abs(frac - 0.5) < 0.4).all(1)]
    h = 1e-6
    _, gx, gy = O.sample(img, xy, interp=interp, dtype=np.float64, with_grad=True)
    fx = (O.sample(img, xy + [h, 0], interp=interp, dtype=np.float64) -
          O.sample(img, xy - [h, 0], interp=interp, dtype=np.float64)) / (2 * h)
    fy = (O.sample(img, xy + [0, h], interp=interp, dtype=np.float64) -
          O.sample(img, xy - [0, h], interp=interp, dtype=np.float64)) / (2 * h)
    np.testing.assert_allclose(gx, fx, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(gy, fy, rtol=1e-5, atol=1e-4)


def test_linear_reference_formula_and_clamp():
    """_bilinear_sample (solver_global_cg.py:481-509) restated in numpy."""
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (16, 16)).astype(np.float32)
    xy = rng.uniform(-3, 19, (400, 2)).astype(np.float32)
    v = O.sample(img, xy, interp="linear", dtype=np.float32)
    x, y = xy[:, 0], xy[:, 1]
    x0 = np.floor(x - np.float32(0.5)).astype(np.int64)
    y0 = np.floor(y - np.float32(0.5)).astype(np.int64)
    fx = (x - np.float32(0.5)) - x0.astype(np.float32)
    fy = (y - np.float32(0.5)) - y0.astype(np.float32)
    c = lambda a: np.clip(a, 0, 15)
    I00, I10 = img[c(y0), c(x0)], img[c(y0), c(x0 + 1)]
    I01, I11 = img[c(y0 + 1), c(x0)], img[c(y0 + 1), c(x0 + 1)]
    one = np.float32(1)
    ref = (one - fx) * (one - fy) * I00 + fx * (one - fy) * I10 + (one - fx) * fy * I01 + fx * fy * I11
    assert np.array_equal(v, ref.astype(np.float32))


def test_cubic_clamp_boundary():
    img = np.arange(36, dtype=np.float64).reshape(6, 6)
    # far outside: value of the nearest corner region
    v = O.sample(img, np.array([[-10.0, -10.0], [50.0, 50.0]]), interp="cubic", dtype=np.float64)
    assert v[0] == img[0, 0] and v[1] == img[-1, -1]
