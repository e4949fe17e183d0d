"""Local Gauss-Newton nodal refinement on the GPU.

Drop-in for ``LocalGaussNewtonSolver`` (reference ``d2ic/solver_local_gn.py:
34-179``): ``config.max_iters`` relaxed Jacobi sweeps of per-node 2x2
Gauss-Newton updates mixing the pixel residuals (through the node->pixel
table of ``PixelAssets``) and the spring term, with central-difference
gradient images of the deformed image (``_local_sweeps`` ``:205-309``,
``jacobi_nodal_step_spring`` ``:312-414``), as HIP kernels behind
``d2ic_lgn_solve``.

``interpolation``: "cubic" (Keys a = -1/2, as in ``GlobalCGSolver``; the
reference delegates cubic to im_jax, which is not available) or
"linear"/"bilinear" (the reference's ``_bilinear_sample``, ``:417-442``).

``layout``: "tiles" (default) runs the sweeps on the global solver's element
tiles (``d2ic_lgn_solve_plan``: per-tile node sums, then a node gathers its
tile sums); "table" walks the reference's node->pixel table
(``d2ic_lgn_solve``).  Both give the same result up to FP32 summation order.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Any

from . import _lib
from .solver_base import SolverBase
from .solver_global_cg import CGPlan, _normalize_interpolation
from .types import Array


@dataclass(frozen=True)
class LocalGNResult:
    """``solver_local_gn.py:27-31``: u (Nn,2), placeholder strain (Nn,3), history
    (max_iters,2) [RMS residual, RMS update] or None."""
    u_nodal: Array
    strain: Array
    history: Array | None = None


class LocalGaussNewtonSolver(SolverBase):
    """Local Gauss-Newton solver for mesh DIC (nodal local updates)."""

    def __init__(self, lam: float = 1e-3, max_step: float = 0.2, omega: float = 0.5, interpolation: str = "cubic",
                 use_map_coordinates: bool | None = None, layout: str = "tiles") -> None:
        if layout not in ("tiles", "table"):
            raise ValueError("layout must be 'tiles' or 'table'")
        self._layout = layout
        self._plan = None
        self._compiled = False
        self._lam = float(lam)
        self._max_step = float(max_step)
        self._omega = float(omega)
        if use_map_coordinates is not None:
            warnings.warn(
                "use_map_coordinates is kept for backward compatibility; use interpolation='cubic' or 'linear'.",
                DeprecationWarning,
                stacklevel=2,
            )
            interpolation = "cubic" if use_map_coordinates else "linear"
        self._interpolation = _normalize_interpolation(interpolation)
        self._dev = None

    def compile(self, assets: Any) -> None:
        """Device copies of the pixel arrays and node tables (``:65-123``)."""
        pix = getattr(assets, "pixel_data", None)
        if pix is None:
            raise ValueError("MeshAssets must provide pixel_data for LocalGaussNewtonSolver.")
        t = _lib.require_gpu()
        f32, i32 = t.float32, t.int32
        d = dict(
            coords=_lib.to_dev(pix.pixel_coords_ref, f32).reshape(-1, 2).contiguous(),
            pnodes=_lib.to_dev(pix.pixel_nodes, i32).reshape(-1, 4).contiguous(),
            shapeN=_lib.to_dev(pix.pixel_shapeN, f32).reshape(-1, 4).contiguous(),
            npi=_lib.to_dev(pix.node_pixel_index, i32).contiguous(),
            npw=_lib.to_dev(pix.node_N_weight, f32).contiguous(),
            npd=_lib.to_dev(pix.node_pixel_degree, i32).reshape(-1).contiguous(),
            nbi=_lib.to_dev(pix.node_neighbor_index, i32).contiguous(),
            nbd=_lib.to_dev(pix.node_neighbor_degree, i32).reshape(-1).contiguous(),
            nbw=_lib.to_dev(pix.node_neighbor_weight, f32).contiguous(),
            nrw=_lib.to_dev(pix.node_reg_weight, f32).reshape(-1).contiguous(),
        )
        nn = int(d["npd"].shape[0])
        for k in ("npi", "npw", "nbi", "nbw"):
            if d[k].dim() != 2 or int(d[k].shape[0]) != nn:
                d[k] = d[k].reshape(nn, -1).contiguous()
        self._dev = d
        self._plan = CGPlan(pix) if self._layout == "tiles" else None
        self._compiled = True

    def solve(self, state: Any, def_image: Array) -> LocalGNResult:
        """``solver_local_gn.py:125-179``."""
        if not self._compiled or self._dev is None:
            raise RuntimeError("LocalGaussNewtonSolver.compile() must be called before solve().")
        if state.assets.pixel_data is None:
            raise ValueError("MeshAssets must provide pixel_data for LocalGaussNewtonSolver.")
        t = _lib.torch()
        d = self._dev
        nn = int(d["npd"].shape[0])
        ref = _lib.to_dev(state.ref_image, t.float32).contiguous()
        de = _lib.to_dev(def_image, t.float32).contiguous()
        if ref.shape != de.shape or ref.dim() != 2:
            raise ValueError("reference and deformed images must be 2-D arrays of the same shape")
        H, W = int(de.shape[0]), int(de.shape[1])
        if state.u0_nodal is not None:
            u = _lib.to_dev(state.u0_nodal, t.float32, (nn, 2)).clone()
        else:
            u = t.zeros((nn, 2), dtype=t.float32, device=_lib.device())
        cfg = state.config
        n_sweeps = int(cfg.max_iters)
        hist = t.full((max(n_sweeps, 0), 2), float("nan"), dtype=t.float32, device=_lib.device())
        hist_p = _lib.ptr(hist) if n_sweeps > 0 else None
        if self._plan is not None:
            _lib.call("d2ic_lgn_solve_plan", self._plan.handle, _lib.ptr(d["nbi"]), _lib.ptr(d["nbd"]),
                      _lib.ptr(d["nbw"]), _lib.ptr(d["nrw"]), int(d["nbi"].shape[1]), _lib.ptr(ref), _lib.ptr(de), H,
                      W, _lib.INTERP[self._interpolation], _lib.ptr(u), n_sweeps, self._lam, self._max_step,
                      float(cfg.reg_strength), self._omega, hist_p, _lib.stream())
        else:
            _lib.call("d2ic_lgn_solve", _lib.ptr(d["coords"]), _lib.ptr(d["pnodes"]), _lib.ptr(d["shapeN"]),
                      int(d["coords"].shape[0]), _lib.ptr(d["npi"]), _lib.ptr(d["npw"]), _lib.ptr(d["npd"]),
                      int(d["npi"].shape[1]), _lib.ptr(d["nbi"]), _lib.ptr(d["nbd"]), _lib.ptr(d["nbw"]),
                      _lib.ptr(d["nrw"]), int(d["nbi"].shape[1]), nn, _lib.ptr(ref), _lib.ptr(de), H, W,
                      _lib.INTERP[self._interpolation], _lib.ptr(u), n_sweeps, self._lam, self._max_step,
                      float(cfg.reg_strength), self._omega, hist_p, _lib.stream())
        strain = t.zeros((nn, 3), dtype=u.dtype, device=u.device)
        save_history = bool(getattr(cfg, "save_history", False))
        return LocalGNResult(u_nodal=u, strain=strain, history=hist if save_history else None)
