"""Synthetic speckle workloads (sum of isotropic Gaussian spots).

Spot parameters come from ``numpy.random.Generator(PCG64(seed))``:
N = round(H*W/20) spots, centres uniform in [0,W) x [0,H), sigma ~ U[1.5, 3],
amplitude ~ U[0.5, 1].  S(x) = sum_k a_k exp(-|x - c_k|^2 / (2 sigma_k^2)) at
pixel centres (j + 0.5, i + 0.5), truncated at a radius of 4 sigma.  Images
are scaled with the reference image's min/max: 255 (S - min S_ref) /
(max S_ref - min S_ref).  A deformed image re-renders the spots at
c_k + u(c_k) and adds N(0, noise_sigma^2) grey-level noise (independent seed).

``render_spots`` here is the numpy renderer (fine up to ~1 Mpx);
``render_spots_device`` (the HIP renderer ``d2ic_render_speckle``) produces
the same sums on the GPU for the large workloads, and
``speckle_sequence_device`` builds whole sequences there (its noise comes from
torch's device generator, so its noisy frames differ from ``speckle_pair``'s).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional, Tuple

import numpy as np


@dataclass(frozen=True)
class Spots:
    cx: np.ndarray
    cy: np.ndarray
    sigma: np.ndarray
    amp: np.ndarray


def speckle_spots(H: int, W: int, seed: int, px_per_spot: float = 20.0) -> Spots:
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    n = int(round(H * W / px_per_spot))
    cx = rng.uniform(0.0, W, n)
    cy = rng.uniform(0.0, H, n)
    sigma = rng.uniform(1.5, 3.0, n)
    amp = rng.uniform(0.5, 1.0, n)
    return Spots(cx=cx, cy=cy, sigma=sigma, amp=amp)


def displaced(spots: Spots, disp_fn: Callable[[np.ndarray, np.ndarray], Tuple[np.ndarray, np.ndarray]]) -> Spots:
    ux, uy = disp_fn(spots.cx, spots.cy)
    return Spots(cx=spots.cx + ux, cy=spots.cy + uy, sigma=spots.sigma, amp=spots.amp)


def render_spots(spots: Spots, H: int, W: int, batch: int = 4096) -> np.ndarray:
    """Float64 sum-of-Gaussians at pixel centres, each spot truncated at 4 sigma."""
    out = np.zeros(H * W, dtype=np.float64)
    R = int(math.ceil(4.0 * 3.0)) + 1
    offs = np.arange(-R, R + 1)
    n = spots.cx.shape[0]
    for s0 in range(0, n, batch):
        cx = spots.cx[s0:s0 + batch, None, None]
        cy = spots.cy[s0:s0 + batch, None, None]
        sg = spots.sigma[s0:s0 + batch, None, None]
        am = spots.amp[s0:s0 + batch, None, None]
        jj = np.floor(cx).astype(np.int64) + offs[None, None, :]
        ii = np.floor(cy).astype(np.int64) + offs[None, :, None]
        dx = (jj + 0.5) - cx
        dy = (ii + 0.5) - cy
        r2 = dx * dx + dy * dy
        val = am * np.exp(-r2 / (2.0 * sg * sg))
        ok = (r2 <= 16.0 * sg * sg) & (jj >= 0) & (jj < W) & (ii >= 0) & (ii < H)
        flat = (ii * W + jj)
        flat = np.broadcast_to(flat, val.shape)
        out += np.bincount(flat[ok], weights=val[ok], minlength=H * W)
    return out.reshape(H, W)


def scale_to_gray(S: np.ndarray, smin: float, smax: float) -> np.ndarray:
    return (255.0 * (S - smin) / (smax - smin)).astype(np.float32)


def speckle_pair(H: int, W: int, seed: int, disp_fn, noise_seed: Optional[int] = None,
                 noise_sigma: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Reference and deformed float32 images (numpy renderer)."""
    spots = speckle_spots(H, W, seed)
    S_ref = render_spots(spots, H, W)
    smin, smax = float(S_ref.min()), float(S_ref.max())
    ref = scale_to_gray(S_ref, smin, smax)
    S_def = render_spots(displaced(spots, disp_fn), H, W)
    dimg = scale_to_gray(S_def, smin, smax)
    if noise_sigma > 0:
        rng = np.random.Generator(np.random.PCG64(seed + 1000003 if noise_seed is None else noise_seed))
        dimg = (dimg + rng.normal(0.0, noise_sigma, size=dimg.shape)).astype(np.float32)
    return ref, dimg


# ----------------------------------------------------------------------------
# displacement families used by the workloads
# ----------------------------------------------------------------------------

def rigid(ux: float, uy: float):
    def f(x, y):
        return np.full_like(x, ux), np.full_like(y, uy)
    return f


def affine_sinusoid(t: float, W: int, H: int, eps: float = 1e-3, amp: float = 1.5, n_frames: int = 20,
                    x0: Optional[float] = None, y0: Optional[float] = None):
    """u(x,y,t) = t [eps (x - x0), -0.3 eps (y - y0)] + amp sin(2 pi x / W) t / n_frames.

    The scalar sinusoid is applied to the y component (a transverse wave)."""
    cx = W / 2.0 if x0 is None else x0
    cy = H / 2.0 if y0 is None else y0

    def f(x, y):
        ux = t * eps * (x - cx)
        uy = t * (-0.3 * eps) * (y - cy) + amp * np.sin(2.0 * np.pi * x / W) * t / n_frames
        return ux, uy
    return f


def affine_field(A, b, amp: float, wavelength: float, cx: float, cy: float):
    """u(x, y) = A [x - cx, y - cy] + b, plus amp sin(2 pi x / wavelength) on u_y."""
    A = np.asarray(A, dtype=np.float64).reshape(2, 2)
    b = np.asarray(b, dtype=np.float64).reshape(2)

    def f(x, y):
        dx, dy = x - cx, y - cy
        ux = A[0, 0] * dx + A[0, 1] * dy + b[0]
        uy = A[1, 0] * dx + A[1, 1] * dy + b[1] + amp * np.sin(2.0 * np.pi * x / wavelength)
        return ux, uy
    return f


def plate_with_hole_mask(H: int, W: int, side: int, radius: float) -> np.ndarray:
    """Centred side x side square minus a centred disc (pixel centres tested)."""
    m = np.zeros((H, W), dtype=bool)
    r0 = (H - side) // 2
    c0 = (W - side) // 2
    m[r0:r0 + side, c0:c0 + side] = True
    yy, xx = np.mgrid[0:H, 0:W]
    d2 = (xx + 0.5 - W / 2.0) ** 2 + (yy + 0.5 - H / 2.0) ** 2
    m &= d2 > radius * radius
    return m


def rect_mask(H: int, W: int, side_h: int, side_w: int) -> np.ndarray:
    m = np.zeros((H, W), dtype=bool)
    r0 = (H - side_h) // 2
    c0 = (W - side_w) // 2
    m[r0:r0 + side_h, c0:c0 + side_w] = True
    return m


# ----------------------------------------------------------------------------
# device renderer (large workloads)
# ----------------------------------------------------------------------------

def spots_array(spots: Spots) -> np.ndarray:
    return np.ascontiguousarray(np.stack([spots.cx, spots.cy, spots.sigma, spots.amp], axis=1), dtype=np.float64)


def render_spots_device(spots: Spots, H: int, W: int):
    """Same sum as ``render_spots`` computed by the HIP renderer; float64 (H, W) device tensor."""
    from . import _lib

    t = _lib.require_gpu()
    arr = t.as_tensor(spots_array(spots), device=_lib.device())
    out = t.empty((H, W), dtype=t.float64, device=_lib.device())
    _lib.call("d2ic_render_speckle", _lib.ptr(arr), int(arr.shape[0]), int(H), int(W), _lib.ptr(out), _lib.stream())
    return out


def speckle_sequence_device(H: int, W: int, seed: int, disp_fns, noise_sigma: float = 1.0):
    """Reference image and one deformed image per displacement function, all
    float32 device tensors (spots, scaling and noise as in ``speckle_pair``)."""
    from . import _lib

    t = _lib.require_gpu()
    spots = speckle_spots(H, W, seed)
    S_ref = render_spots_device(spots, H, W)
    smin, smax = float(S_ref.min()), float(S_ref.max())
    ref = (255.0 * (S_ref - smin) / (smax - smin)).to(t.float32)
    frames = []
    gen = t.Generator(device=_lib.device())
    for k, fn in enumerate(disp_fns):
        S = render_spots_device(displaced(spots, fn), H, W)
        img = (255.0 * (S - smin) / (smax - smin)).to(t.float32)
        if noise_sigma > 0:
            gen.manual_seed(int(seed) * 7919 + 1000003 + k)
            img = img + noise_sigma * t.randn(img.shape, generator=gen, device=_lib.device(), dtype=t.float32)
        frames.append(img.contiguous())
        del S
    return ref.contiguous(), frames
