"""CPU oracle for the D2IC global FE-DIC hot path -- TEST INFRASTRUCTURE ONLY.

This module wraps ``oracle/_build/liboracle.so`` (built from ``d2ic_oracle.c``
by ``oracle/Makefile``).  It is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and never as the thing measured or shipped.  The product package
``d2ic_amd`` does not import it.

Every function takes ``dtype`` = ``np.float32`` (the reference's default JAX
FP32 run) or ``np.float64`` (its x64 run).

What pins each function (tests named in brackets):

* ``pixel_candidates`` / ``pixel_to_element`` -- the reference's exact integer
  fixture (``tests/test_pixel_assets.py:6-36``) [test_oracle_assets], brute
  force vs. binned search on random convex quads, and the closed form on
  structured grids (element = cell index) [test_oracle_assets].
* ``shape_functions`` -- closed form on axis-aligned rectangles
  (xi = 2(x - xc)/w), partition of unity, non-negativity, and
  x(xi, eta) = p on distorted quads [test_oracle_assets].
* ``node_pixel_dense`` -- ascending-pixel closed form on structured grids,
  degree sum = 4 Np, and a numpy argsort restatement [test_oracle_assets].
* ``node_neighbor_dense`` -- numpy unique/lexsort restatement and the exact
  weights 1/(16 + 1e-6), 1/(16 sqrt 2 + 1e-6) [test_oracle_assets].
* ``sample`` (linear) -- the in-repo formula: exact at pixel centres, exact
  on affine images, clamp at the border [test_oracle_sampling].
  ``sample`` (cubic) -- PARITY UNPINNED against the reference: the reference
  calls im_jax.flat_nd_cubic_interpolate, which is not available.  The Keys
  a=-1/2 kernel is pinned only by its own invariants: interpolation at
  centres, reproduction of constants/ramps/quadratics, clamp boundary, and
  derivative vs. central differences [test_oracle_sampling].
* ``objective`` -- closed forms (zero residual at the true translation of an
  affine image; E_reg = 0 for uniform u; reg gradient = L_W u) and gradient
  vs. central finite differences in FP64 [test_oracle_objective].
* ``cg_solve`` -- recovers a known rigid translation on synthetic speckle,
  n_iters/history semantics (stop before update, NaN padding, alpha = 1 on
  the first accepted trial) and FP32 vs FP64 agreement [test_oracle_objective];
  the NLCG loop past iteration 0 against a numpy restatement [test_oracle_nlcg].
* ``zncc`` -- the reference's known-shift fixture
  (``tests/test_solver_translation_zncc.py:8-33``) and a numpy brute force
  [test_oracle_zncc].
* ``idw`` -- closed form on tiny inputs and a numpy restatement [test_oracle_strain_idw].
* ``strain`` -- the reference's closed form (``tests/test_strain_smoke.py:10-40``)
  and exact E for affine fields [test_oracle_strain_idw].
* ``image_gradient`` -- exact on ramps, zero edges [test_oracle_local_gn].
  ``local_gn_step`` -- closed forms: one step recovers a displacement when the
  residual is linear in it, lam shrinks an isotropic step by 1/(1 + 2 lam),
  max_step clips the norm, omega scales the update, a node with only springs
  moves to the weighted mean of its neighbours, a node with neither pixels nor
  springs stays [test_oracle_local_gn].  ``local_gn_sweeps`` -- recovers a
  sub-pixel translation of synthetic speckle with a decreasing residual
  [test_oracle_local_gn].
* ``bilinear_nan`` / ``discrepancy_map`` -- exact at pixel centres, exact on
  affine images (residual = a ux + b uy for a constant u), the validity edge
  (the last image row/column is outside, the sample just inside is not), NaN
  RMS as soon as one sample is invalid, duplicated image pixels averaged and
  poisoned by an invalid one, and no pixels (all NaN, RMS 0)
  [test_oracle_discrepancy].
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liboracle.so"
if os.environ.get("D2IC_ORACLE_SANITIZED") == "1":
    # AddressSanitizer + UBSan build of the same source (``make sanitized``);
    # the process must have loaded libasan first
    _LIB_PATH = _HERE / "_build" / "liboracle_san.so"
_lib = None

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_int = ctypes.c_int
_vp = ctypes.c_void_p
_dbl = ctypes.c_double


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc)."""
    if force or not _LIB_PATH.exists():
        target = "sanitized" if _LIB_PATH.name == "liboracle_san.so" else "all"
        subprocess.run(["make", "-s", "-C", str(_HERE), target], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        for sfx, real in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            getattr(_lib, f"orc_objective_{sfx}").restype = _dbl
            getattr(_lib, f"orc_pixel_candidates_{sfx}").restype = _i64
            getattr(_lib, f"orc_first_element_{sfx}").restype = _i32
    return _lib


def _sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise ValueError(f"unsupported dtype {dtype}")


def _real(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def num_threads() -> int:
    return int(lib().orc_num_threads_f32())


# --------------------------------------------------------------------------
# assets
# --------------------------------------------------------------------------

def pixel_to_element(pixel_xy, nodes, elements, dtype=np.float32, use_bins=True):
    s = _sfx(dtype)
    xy = _c(pixel_xy, dtype).reshape(-1, 2)
    nd = _c(nodes, dtype).reshape(-1, 2)
    el = _c(elements, np.int32).reshape(-1, 4)
    n = xy.shape[0]
    elts = np.empty(n, np.int32)
    pn = np.empty((n, 4), np.int32)
    rc = getattr(lib(), f"orc_pixel_to_element_{s}")(_p(xy), _i64(n), _p(nd), _p(el), _i64(el.shape[0]),
                                                        _int(int(use_bins)), _p(elts), _p(pn))
    if rc != 0:
        raise RuntimeError("orc_pixel_to_element failed")
    return elts, pn


def shape_functions(pixel_xy, pixel_nodes, nodes, dtype=np.float32):
    s = _sfx(dtype)
    xy = _c(pixel_xy, dtype).reshape(-1, 2)
    pn = _c(pixel_nodes, np.int32).reshape(-1, 4)
    nd = _c(nodes, dtype).reshape(-1, 2)
    n = xy.shape[0]
    N = np.empty((n, 4), dtype)
    xe = np.empty((n, 2), dtype)
    getattr(lib(), f"orc_shape_functions_{s}")(_p(xy), _p(pn), _i64(n), _p(nd), _p(N), _p(xe))
    return N, xe


def pixel_candidates(nodes, elements, H, W, roi_mask=None, pad=2.0, dtype=np.float32, use_bins=True):
    """Candidate sampling + ROI filter + element assignment + drop
    (pixel_assets.py:58-79).  Returns coords, pixel_nodes, pixel_elts, roi_flat."""
    s = _sfx(dtype)
    nd = _c(nodes, dtype).reshape(-1, 2)
    el = _c(elements, np.int32).reshape(-1, 4)
    cap = int(H) * int(W)
    coords = np.empty((cap, 2), dtype)
    pn = np.empty((cap, 4), np.int32)
    pe = np.empty(cap, np.int32)
    rf = np.empty(cap, np.int64)
    roi = None
    if roi_mask is not None:
        roi = np.ascontiguousarray(np.asarray(roi_mask, dtype=bool).astype(np.uint8))
    n = getattr(lib(), f"orc_pixel_candidates_{s}")(_p(nd), _i64(nd.shape[0]), _p(el), _i64(el.shape[0]),
                                                      _int(int(H)), _int(int(W)), _p(roi), _dbl(float(pad)),
                                                      _int(int(use_bins)), _i64(cap), _p(coords), _p(pn), _p(pe),
                                                      _p(rf))
    if n < 0:
        raise RuntimeError("orc_pixel_candidates failed")
    return coords[:n].copy(), pn[:n].copy(), pe[:n].copy(), rf[:n].copy()


def node_pixel_dense(pixel_nodes, shapeN, n_nodes, dtype=np.float32):
    s = _sfx(dtype)
    pn = _c(pixel_nodes, np.int32).reshape(-1, 4)
    N = _c(shapeN, dtype).reshape(-1, 4)
    Np = pn.shape[0]
    deg = np.empty(n_nodes, np.int32)
    md = _i64(0)
    f = getattr(lib(), f"orc_node_pixel_dense_{s}")
    f(_p(pn), _p(N), _i64(Np), _i64(n_nodes), _i64(0), _p(deg), ctypes.byref(md), None, None)
    maxdeg = int(md.value)
    idx = np.empty((n_nodes, maxdeg), np.int32)
    w = np.empty((n_nodes, maxdeg), dtype)
    if maxdeg > 0:
        f(_p(pn), _p(N), _i64(Np), _i64(n_nodes), _i64(maxdeg), _p(deg), ctypes.byref(md), _p(idx), _p(w))
    return idx, w, deg


def node_neighbor_dense(elements, nodes, n_nodes, dtype=np.float32):
    s = _sfx(dtype)
    el = _c(elements, np.int32).reshape(-1, 4)
    nd = _c(nodes, dtype).reshape(-1, 2)
    cap = 64
    while True:
        idx = -np.ones((n_nodes, cap), np.int32)
        w = np.zeros((n_nodes, cap), dtype)
        deg = np.empty(n_nodes, np.int32)
        md = _i64(0)
        rc = getattr(lib(), f"orc_node_neighbor_dense_{s}")(_p(el), _i64(el.shape[0]), _p(nd), _i64(n_nodes),
                                                              _i64(cap), _p(idx), _p(deg), _p(w), ctypes.byref(md))
        if rc == 0:
            m = int(md.value)
            return idx[:, :m].copy(), deg, w[:, :m].copy()
        cap *= 2


def pixel_assets(nodes, elements, image_shape, roi_mask=None, pad=2.0, dtype=np.float32):
    """Full build_pixel_assets restatement (pixel_assets.py:21-112) as a dict."""
    H, W = image_shape
    nodes = np.asarray(nodes, dtype)
    coords, pn, pe, rf = pixel_candidates(nodes, elements, H, W, roi_mask=roi_mask, pad=pad, dtype=dtype)
    N, _ = shape_functions(coords, pn, nodes, dtype=dtype)
    nn = nodes.shape[0]
    nb_idx, nb_deg, nb_w = node_neighbor_dense(elements, nodes, nn, dtype=dtype)
    npi, npw, npd = node_pixel_dense(pn, N, nn, dtype=dtype)
    return dict(
        pixel_coords_ref=coords, pixel_nodes=pn, pixel_shapeN=N, pixel_elts=pe,
        node_neighbor_index=nb_idx, node_neighbor_degree=nb_deg, node_neighbor_weight=nb_w,
        node_reg_weight=np.ones(nn, dtype),
        node_pixel_index=npi, node_N_weight=npw, node_pixel_degree=npd,
        roi_mask_flat=rf, image_shape=(int(H), int(W)),
    )


# --------------------------------------------------------------------------
# sampling / objective / NLCG
# --------------------------------------------------------------------------

_INTERP = {"cubic": 0, "linear": 1, "bilinear": 1}


def sample(image, xy, interp="linear", dtype=np.float32, with_grad=False):
    s = _sfx(dtype)
    img = _c(image, dtype)
    H, W = img.shape
    pts = _c(xy, dtype).reshape(-1, 2)
    n = pts.shape[0]
    v = np.empty(n, dtype)
    gx = np.empty(n, dtype) if with_grad else None
    gy = np.empty(n, dtype) if with_grad else None
    getattr(lib(), f"orc_sample_{s}")(_p(img), _i64(H), _i64(W), _p(pts), _i64(n), _int(_INTERP[interp]),
                                      _p(v), _p(gx), _p(gy))
    return (v, gx, gy) if with_grad else v


def _asset_args(a, dtype):
    coords = _c(a["pixel_coords_ref"], dtype)
    pn = _c(a["pixel_nodes"], np.int32)
    N = _c(a["pixel_shapeN"], dtype)
    npi = _c(a["node_pixel_index"], np.int32)
    npw = _c(a["node_N_weight"], dtype)
    npd = _c(a["node_pixel_degree"], np.int32)
    nbi = _c(a["node_neighbor_index"], np.int32)
    nbd = _c(a["node_neighbor_degree"], np.int32)
    nbw = _c(a["node_neighbor_weight"], dtype)
    rw = _c(a["node_reg_weight"], dtype)
    keep = (coords, pn, N, npi, npw, npd, nbi, nbd, nbw, rw)
    args = [_p(coords), _p(pn), _p(N), _i64(coords.shape[0]), _p(npi), _p(npw), _p(npd),
            _i64(npi.shape[1] if npi.ndim == 2 else 0), _p(nbi), _p(nbd), _p(nbw), _p(rw),
            _i64(nbi.shape[1] if nbi.ndim == 2 else 0), _i64(rw.shape[0])]
    return keep, args


def objective(u, ref_image, def_image, assets, alpha_reg=0.0, interp="cubic", dtype=np.float32, with_grad=True):
    s = _sfx(dtype)
    R = _real(dtype)
    uu = _c(u, dtype).reshape(-1, 2)
    f = _c(ref_image, dtype)
    g = _c(def_image, dtype)
    H, W = f.shape
    keep, args = _asset_args(assets, dtype)
    grad = np.empty_like(uu) if with_grad else None
    J = getattr(lib(), f"orc_objective_{s}")(_p(uu), _p(f), _p(g), _i64(H), _i64(W), _int(_INTERP[interp]),
                                             R(alpha_reg), *args, _p(grad))
    return (float(J), grad) if with_grad else float(J)


def cg_solve(u0, ref_image, def_image, assets, alpha_reg=0.0, max_iter=200, tol=1e-3, interp="cubic",
             dtype=np.float32):
    """_cg_solve restatement.  Returns (u, hist (max_iter,3), n_iters)."""
    s = _sfx(dtype)
    R = _real(dtype)
    u = np.array(u0, dtype=dtype).reshape(-1, 2).copy()
    f = _c(ref_image, dtype)
    g = _c(def_image, dtype)
    H, W = f.shape
    keep, args = _asset_args(assets, dtype)
    hist = np.empty((int(max_iter), 3), dtype)
    nit = ctypes.c_int(0)
    getattr(lib(), f"orc_cg_solve_{s}")(_p(u), _p(f), _p(g), _i64(H), _i64(W), _int(_INTERP[interp]), R(alpha_reg),
                                        _int(int(max_iter)), R(tol), *args, _p(hist), ctypes.byref(nit))
    return u, hist, int(nit.value)


# --------------------------------------------------------------------------
# ZNCC / IDW / strain
# --------------------------------------------------------------------------

def zncc(centers, ref_image, def_image, win, search, dtype=np.float32):
    s = _sfx(dtype)
    c = _c(centers, dtype).reshape(-1, 2)
    r = _c(ref_image, dtype)
    d = _c(def_image, dtype)
    u = np.empty((c.shape[0], 2), dtype)
    sc = np.empty(c.shape[0], dtype)
    getattr(lib(), f"orc_zncc_{s}")(_p(c), _i64(c.shape[0]), _p(r), _i64(r.shape[0]), _i64(r.shape[1]), _p(d),
                                    _i64(d.shape[0]), _i64(d.shape[1]), _int(int(win)), _int(int(search)), _p(u),
                                    _p(sc))
    return u, sc


def idw(nodes, centers, disp, dtype=np.float32):
    s = _sfx(dtype)
    nd = _c(nodes, dtype).reshape(-1, 2)
    c = _c(centers, dtype).reshape(-1, 2)
    d = _c(disp, dtype).reshape(-1, 2)
    valid = ~np.isnan(d).any(axis=1)
    c = np.ascontiguousarray(c[valid])
    d = np.ascontiguousarray(d[valid])
    out = np.zeros_like(nd)
    if c.shape[0] == 0:
        return out
    getattr(lib(), f"orc_idw_{s}")(_p(nd), _i64(nd.shape[0]), _p(c), _p(d), _i64(c.shape[0]), _p(out))
    return out


def strain(u, nodes, nb_idx, nb_deg, gauge_length=0.0, eps=1e-8, dtype=np.float32):
    """Returns (F (Nn,2,2), E (Nn,2,2), voigt (Nn,3))."""
    s = _sfx(dtype)
    R = _real(dtype)
    uu = _c(u, dtype).reshape(-1, 2)
    X = _c(nodes, dtype).reshape(-1, 2)
    bi = _c(nb_idx, np.int32)
    bd = _c(nb_deg, np.int32)
    n = X.shape[0]
    F = np.empty((n, 2, 2), dtype)
    E = np.empty((n, 2, 2), dtype)
    V = np.empty((n, 3), dtype)
    getattr(lib(), f"orc_strain_{s}")(_p(uu), _p(X), _i64(n), _p(bi), _p(bd), _i64(bi.shape[1]), R(gauge_length),
                                      R(eps), _p(F), _p(E), _p(V))
    return F, E, V


# --------------------------------------------------------------------------
# local Gauss-Newton refinement (numpy restatement over the C sampler)
# --------------------------------------------------------------------------

def image_gradient(im, dtype=np.float32):
    """Central differences with zero edges (solver_local_gn.py:186-202)."""
    im = np.asarray(im, dtype)
    gx = np.zeros_like(im)
    gy = np.zeros_like(im)
    half = dtype(0.5)
    gx[:, 1:-1] = half * (im[:, 2:] - im[:, :-2])
    gy[1:-1, :] = half * (im[2:, :] - im[:-2, :])
    return gx, gy


def local_gn_step(u, r, gx, gy, assets, lam=0.1, max_step=0.2, alpha_reg=0.0, omega=0.5, dtype=np.float32):
    """jacobi_nodal_step_spring (solver_local_gn.py:312-414) for per-pixel r and
    sampled gradient images gx, gy; returns the relaxed update of u (Nn, 2)."""
    R = dtype
    u = np.asarray(u, R).reshape(-1, 2)
    r = np.asarray(r, R)
    gx = np.asarray(gx, R)
    gy = np.asarray(gy, R)
    npi = np.asarray(assets["node_pixel_index"])
    npw = np.asarray(assets["node_N_weight"], R)
    deg_pix = np.asarray(assets["node_pixel_degree"]).reshape(-1)
    nbi = np.asarray(assets["node_neighbor_index"])
    nbw = np.asarray(assets["node_neighbor_weight"], R)
    deg_nb = np.asarray(assets["node_neighbor_degree"]).reshape(-1)
    nrw = np.asarray(assets["node_reg_weight"], R).reshape(-1)
    nn = u.shape[0]
    mdp = npi.shape[1] if npi.ndim == 2 else 0
    mdn = nbi.shape[1] if nbi.ndim == 2 else 0
    pm = np.arange(mdp)[None, :] < deg_pix[:, None]
    idx = np.where(pm, npi, 0) if mdp else np.zeros((nn, 0), np.int64)
    Ni = np.where(pm, npw, R(0)) if mdp else np.zeros((nn, 0), R)
    Jx = gx[idx] * Ni
    Jy = gy[idx] * Ni
    ri = r[idx]
    g_img = np.stack([(ri * Jx).sum(1), (ri * Jy).sum(1)], 1).astype(R)
    H00 = (Jx * Jx).sum(1)
    H01 = (Jx * Jy).sum(1)
    H11 = (Jy * Jy).sum(1)
    lam_loc = R(lam) * (H00 + H11 + R(1e-12))
    nm = np.arange(mdn)[None, :] < deg_nb[:, None]
    ids = np.where(nm, nbi, 0) if mdn else np.zeros((nn, 0), np.int64)
    w = np.where(nm, nbw, R(0)) if mdn else np.zeros((nn, 0), R)
    un = np.where(nm[..., None], u[ids], R(0)) if mdn else np.zeros((nn, 0, 2), R)
    diff = u[:, None, :] - un
    alpha_loc = R(alpha_reg) * nrw
    g_reg = alpha_loc[:, None] * (w[..., None] * diff).sum(1)
    sumw = w.sum(1)
    H = np.zeros((nn, 2, 2), R)
    H[:, 0, 0] = H00 + lam_loc + alpha_loc * sumw + R(1e-8)
    H[:, 0, 1] = H01
    H[:, 1, 0] = H01
    H[:, 1, 1] = H11 + lam_loc + alpha_loc * sumw + R(1e-8)
    g = (g_img + g_reg).astype(R)
    delta = -np.linalg.solve(H, g[..., None])[..., 0].astype(R)
    has = (deg_pix > 0) | ((R(alpha_reg) != 0) & (deg_nb > 0))
    delta = np.where(has[:, None], delta, R(0))
    nrm = np.sqrt((delta * delta).sum(1))
    factor = np.minimum(R(1), R(max_step) / (nrm + R(1e-12)))
    delta = delta * factor[:, None]
    return (u + R(omega) * delta).astype(R)


def local_gn_sweeps(u0, ref_image, def_image, assets, lam=1e-3, max_step=0.2, alpha_reg=0.0, omega=0.5, n_sweeps=3,
                    interp="cubic", dtype=np.float32):
    """_local_sweeps with _pixel_state (solver_local_gn.py:205-309): returns
    (u, history (n_sweeps, 2) = [RMS residual, RMS update])."""
    R = dtype
    gx2, gy2 = image_gradient(def_image, R)
    coords = np.asarray(assets["pixel_coords_ref"], R).reshape(-1, 2)
    pn = np.asarray(assets["pixel_nodes"]).reshape(-1, 4)
    N = np.asarray(assets["pixel_shapeN"], R).reshape(-1, 4)
    I1 = sample(ref_image, coords, interp, dtype=R)
    u = np.asarray(u0, R).reshape(-1, 2).copy()
    hist = np.full((int(n_sweeps), 2), np.nan, R)
    for k in range(int(n_sweeps)):
        upix = (N[..., None] * u[pn]).sum(1)
        xd = (coords + upix).astype(R)
        r = (sample(def_image, xd, interp, dtype=R) - I1).astype(R)
        gxd = sample(gx2, xd, interp, dtype=R)
        gyd = sample(gy2, xd, interp, dtype=R)
        un = local_gn_step(u, r, gxd, gyd, assets, lam=lam, max_step=max_step, alpha_reg=alpha_reg, omega=omega,
                           dtype=R)
        step = un - u
        hist[k, 0] = np.sqrt(np.mean(r.astype(np.float64) ** 2))
        hist[k, 1] = np.sqrt(np.mean(step.astype(np.float64) ** 2))
        u = un
    return u, hist


# --------------------------------------------------------------------------
# discrepancy map (numpy restatement, FP32 in the reference's operation order)
# --------------------------------------------------------------------------

def bilinear_nan(image, coords, dtype=np.float32):
    """bilinear_sample_jax (discrepancy.py:51-83): samples at (x, y) - 0.5,
    NaN unless 0 <= floor(y - .5) < H - 1 and 0 <= floor(x - .5) < W - 1."""
    R = dtype
    image = np.asarray(image, R)
    coords = np.asarray(coords, R).reshape(-1, 2)
    H, W = image.shape
    xc = coords[:, 0] - R(0.5)
    yc = coords[:, 1] - R(0.5)
    j0 = np.floor(xc).astype(np.int64)
    i0 = np.floor(yc).astype(np.int64)
    fx = xc - j0.astype(R)
    fy = yc - i0.astype(R)
    valid = (i0 >= 0) & (i0 < H - 1) & (j0 >= 0) & (j0 < W - 1)
    i0v = np.clip(i0, 0, max(H - 2, 0))
    j0v = np.clip(j0, 0, max(W - 2, 0))
    v00 = image[i0v, j0v]
    v01 = image[i0v, j0v + 1]
    v10 = image[i0v + 1, j0v]
    v11 = image[i0v + 1, j0v + 1]
    one = R(1.0)
    vals = v00 * (one - fx) * (one - fy) + v01 * fx * (one - fy) + v10 * (one - fx) * fy + v11 * fx * fy
    return np.where(valid, vals, R(np.nan)).astype(R)


def discrepancy_map(ref_image, def_image, u, assets, dtype=np.float32):
    """compute_discrepancy_map_ref (discrepancy.py:86-127): returns
    (map (H, W), rms).  r = I_def(x + u(x)) - I_ref(x); the map is
    bincount(roi_flat, r * valid) / max(bincount(valid), 1), NaN where no
    valid pixel landed; rms = sqrt(sum((r * valid)^2) / max(sum(valid), 1)),
    which is NaN once any r is NaN (NaN * 0)."""
    R = dtype
    H, W = (int(v) for v in assets["image_shape"])
    coords = np.asarray(assets["pixel_coords_ref"], R).reshape(-1, 2)
    pn = np.asarray(assets["pixel_nodes"]).reshape(-1, 4)
    N = np.asarray(assets["pixel_shapeN"], R).reshape(-1, 4)
    rf = np.asarray(assets["roi_mask_flat"]).reshape(-1).astype(np.int64)
    u = np.asarray(u, R).reshape(-1, 2)
    prod = N[..., None] * u[pn]
    disp = prod[:, 0]
    for k in range(1, 4):
        disp = disp + prod[:, k]
    i_ref = bilinear_nan(ref_image, coords, R)
    i_def = bilinear_nan(def_image, coords + disp, R)
    with np.errstate(invalid="ignore"):
        res = (i_def - i_ref).astype(R)
        valid = np.isfinite(res)
        vf = valid.astype(R)
        w = res * vf
        denom = max(float(vf.sum(dtype=np.float64)), 1.0)
        rms = float(np.sqrt((w.astype(np.float64) ** 2).sum() / denom))
        acc = np.zeros(H * W, R)
        cnt = np.zeros(H * W, R)
        np.add.at(acc, rf, w)
        np.add.at(cnt, rf, vf)
        out = np.where(cnt > 0, acc / np.maximum(cnt, R(1)), R(np.nan)).astype(R)
    return out.reshape(H, W), rms
