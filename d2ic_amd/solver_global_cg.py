"""Global NLCG solver for mesh DIC on the GPU.

Drop-in for ``GlobalCGSolver`` (reference ``d2ic/solver_global_cg.py:35-198``).
It minimises J(u) = 1/2 mean_p (g(x_p + u(x_p)) - f(x_p))^2 + reg_strength *
E_spring(u) with nonlinear CG (Polak-Ribiere+, restart on non-descent,
10-trial Armijo backtracking with c = 1e-4, stop when |grad J|_2 < tol before
the update) -- the reference's ``_cg_solve`` (``:326-469``) -- in one
persistent HIP kernel per solve (``d2ic_cg_solve``).

``interpolation``: "cubic" (Keys a = -1/2 cubic convolution with clamp-to-edge
boundaries; the reference delegates cubic to im_jax, whose kernel is not
available) or "linear"/"bilinear" (the reference's in-repo bilinear sampler,
``_bilinear_sample``, ``:481-509``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import json
import os
import sys
import warnings
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _lib
from .mesh_assets import MeshAssets, PixelAssets
from .solver_base import SolverBase
from .types import Array


@dataclass(frozen=True)
class GlobalCGResult:
    """``solver_global_cg.py:27-32``: u (Nn,2), placeholder strain (Nn,2),
    history (max_iter,3) [J, |grad|, alpha] or None, iteration count."""
    u_nodal: Array
    strain: Array
    history: Array | None = None
    n_iters: int | None = None


class PendingSolve:
    """A queued solve (GlobalCGSolver.launch).  ``result`` holds u (filled in
    place by the queued kernels) and the history buffer; ``finish()`` waits
    for the iteration count and returns the completed GlobalCGResult.  Keeps
    the deformed image alive until then."""

    def __init__(self, solver, result, hist, counters, image):
        self._solver = solver
        self.result = result
        self.hist = hist
        self.counters = counters
        self._image = image
        self._done = False

    def finish(self) -> GlobalCGResult:
        if not self._done:
            self._solver._finish(self)
            self._done = True
            self._image = None
        return self.result


def _normalize_interpolation(interpolation: str) -> str:
    """``solver_global_cg.py:472-478``."""
    value = str(interpolation).lower()
    if value == "bilinear":
        value = "linear"
    if value not in ("cubic", "linear"):
        raise ValueError(f"Unsupported interpolation: {interpolation}")
    return value


def _autotune_cache_path(cache):
    """File of the persisted grid-autotune results (None: no cache)."""
    if cache is False:
        return None
    if cache is None:
        cache = os.environ.get("D2IC_AUTOTUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "d2ic_amd", "grid_autotune.json")
    return str(cache)


def _autotune_cache_load(path) -> dict:
    try:
        with open(path, "r", encoding="utf-8") as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _autotune_cache_store(path, key, entry) -> None:
    """Add one entry (read-modify-write, replaced atomically; a failure to
    write leaves the tuning result in effect and is not an error)."""
    try:
        data = _autotune_cache_load(path)
        data[key] = entry
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w", encoding="utf-8") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


class CGPlan:
    """Owner of a device plan (``d2ic_cg_plan``) built from PixelAssets.

    ``tile_px`` (64..256) and ``strip_w`` (4..24) choose the tiling: tiles
    close at ``tile_px`` pixels, and elements wider than 24 px are cut into
    column strips of at most ``strip_w`` pixels.  The defaults (256, 16) are
    the plan ``d2ic_cg_plan_create`` builds."""

    DEFAULT_TILING = (256, 16)

    def __init__(self, pix: PixelAssets, tile_px: int = 256, strip_w: int = 16):
        t = _lib.require_gpu()
        self._keep = []

        def dev(a, dtype):
            x = _lib.to_dev(a, dtype)
            self._keep.append(x)
            return x

        coords = dev(pix.pixel_coords_ref, t.float32)
        pnodes = dev(pix.pixel_nodes, t.int32)
        shapeN = dev(pix.pixel_shapeN, t.float32)
        pelts = dev(pix.pixel_elements, t.int32) if getattr(pix, "pixel_elements", None) is not None else None
        nb_idx = dev(pix.node_neighbor_index, t.int32)
        nb_deg = dev(pix.node_neighbor_degree, t.int32)
        nb_w = dev(pix.node_neighbor_weight, t.float32)
        reg_w = dev(pix.node_reg_weight, t.float32)
        self.n_pixels = int(coords.shape[0])
        self.n_nodes = int(nb_deg.shape[0])
        # nodes that move with the images (pixels or spring neighbours); the others keep u0
        self.node_active = (dev(pix.node_pixel_degree, t.int32) > 0) | (nb_deg > 0)
        md = int(nb_idx.shape[1]) if nb_idx.dim() == 2 else 0
        h = ctypes.c_void_p(0)
        self.tiling = (int(tile_px), int(strip_w))
        _lib.call("d2ic_cg_plan_create_tiled", _lib.ptr(coords), _lib.ptr(pnodes), _lib.ptr(shapeN), _lib.ptr(pelts),
                  self.n_pixels, self.n_nodes, _lib.ptr(nb_idx), _lib.ptr(nb_deg), _lib.ptr(nb_w), _lib.ptr(reg_w),
                  md, self.tiling[0], self.tiling[1], ctypes.byref(h), _lib.stream())
        self.handle = h
        self._ref_key = None
        self.grid_blocks = 0
        self._keep = []  # the plan copied what it needs

    def info(self):
        a, b, c, d = (ctypes.c_int64(0) for _ in range(4))
        _lib.call("d2ic_cg_plan_info", self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                  ctypes.byref(d))
        return dict(n_pixels=a.value, n_nodes=b.value, n_tiles=c.value, device_bytes=d.value)

    def kernel_ms(self) -> float:
        """Duration of the last persistent CG kernel launch this plan led (waits for it)."""
        ms = ctypes.c_float(0.0)
        _lib.call("d2ic_cg_kernel_ms", self.handle, ctypes.byref(ms))
        return float(ms.value)

    def set_grid(self, blocks: int) -> None:
        """Persistent grid size (blocks) for single solves on this plan; 0
        restores the work-based default.  Only speed depends on it."""
        _lib.call("d2ic_cg_plan_set_grid", self.handle, int(blocks))
        self.grid_blocks = int(blocks)

    def autotune(self, def_image, alpha_reg: float = 0.1, max_iter: int = 8, candidates=None,
                 repeats: int = 3, cache=None) -> dict:
        """Choose the persistent grid size by timing short solves.

        Each candidate grid (blocks, multiples of 8 up to one block per CU,
        and 0 = the work-based default) runs ``repeats`` solves of
        ``max_iter`` iterations from u = 0 with tol = 0, so every candidate
        does the same passes; the kernel time is read from the HIP events
        around the launch and the fastest median wins and is kept
        (``set_grid``).  The reference image must be set.  The grid is the
        one launch parameter that leaves results bitwise unchanged (every
        reduction is an order-independent fixed-point sum); the tiling
        changes rounding and is chosen with the plan
        (``GlobalCGSolver.autotune_tiling``).

        ``cache``: path of a JSON file keyed on (device, interpolation, image
        shape, pixels, tiling, tiles, nodes, candidates, iterations); None uses
        ``$D2IC_AUTOTUNE_CACHE`` or ``~/.cache/d2ic_amd/grid_autotune.json``,
        False disables it.  A cached entry is applied without timing.
        Returns ``{"blocks": best, "ms": {blocks: median ms}, "default_ms":
        ms of grid 0, "cached": bool}``."""
        t = _lib.torch()
        if self._ref_key is None:
            raise RuntimeError("set_reference() must be called before autotune()")
        g = _lib.to_dev(def_image, t.float32)
        if candidates is None:
            cu = int(t.cuda.get_device_properties(_lib.device()).multi_processor_count)
            candidates = sorted({max(8, (cu * k // 8 + 7) // 8 * 8) for k in (2, 4, 6, 8)})
        cands = sorted({0, *(int(b) for b in candidates)})
        path = _autotune_cache_path(cache)
        key = None
        if path is not None:
            info = self.info()
            key = "|".join(str(v) for v in (
                t.cuda.get_device_name(_lib.device()), self._ref_key[2], "x".join(map(str, self._ref_key[1])),
                info["n_pixels"], "{}/{}".format(*self.tiling), info["n_tiles"], info["n_nodes"],
                ",".join(map(str, cands)), int(max_iter)))
            hit = _autotune_cache_load(path).get(key)
            if hit is not None and int(hit["blocks"]) in cands:
                self.set_grid(int(hit["blocks"]))
                ms = {int(k): float(v) for k, v in hit["ms"].items()}
                return {"blocks": int(hit["blocks"]), "ms": ms, "default_ms": ms.get(0), "cached": True}
        times = {}
        for b in cands:
            self.set_grid(int(b))
            times[int(b)] = self.time_solves(g, alpha_reg, max_iter, repeats)
        best = min(times, key=times.get)
        self.set_grid(best)
        if path is not None:
            _autotune_cache_store(path, key, {"blocks": best, "ms": {str(k): v for k, v in times.items()}})
        return {"blocks": best, "ms": times, "default_ms": times[0], "cached": False}

    def time_solves(self, def_image, alpha_reg: float = 0.1, max_iter: int = 8, repeats: int = 3,
                    u0=None) -> float:
        """Median kernel time (ms) of ``repeats`` solves of ``max_iter``
        iterations with tol = 0 from ``u0`` (default zeros), after one
        dropped warm-up solve.  The reference image must be set."""
        t = _lib.torch()
        g = _lib.to_dev(def_image, t.float32)
        start = (t.zeros((self.n_nodes, 2), dtype=t.float32, device=_lib.device()) if u0 is None
                 else _lib.to_dev(u0, t.float32, (self.n_nodes, 2)))
        u = t.empty_like(start)
        cnt = t.zeros((2,), dtype=t.int32, device=_lib.device())
        ms = []
        for _ in range(int(repeats) + 1):
            u.copy_(start)
            _lib.call("d2ic_cg_solve", self.handle, _lib.ptr(g), int(g.shape[0]), int(g.shape[1]), _lib.ptr(u),
                      float(alpha_reg), int(max_iter), 0.0, None, _lib.ptr(cnt), _lib.ptr(cnt) + 4, _lib.stream())
            ms.append(self.kernel_ms())
        if int(cnt[0].item()) < 0:
            raise RuntimeError("d2ic_cg_solve did not finish (a grid barrier timed out)")
        return float(np.median(ms[1:]))

    def set_reference(self, ref_image, interpolation: str):
        t = _lib.torch()
        ref = _lib.to_dev(ref_image, t.float32)
        key = (ref.data_ptr(), tuple(ref.shape), interpolation, getattr(ref, "_version", 0))
        if key == self._ref_key:
            return ref
        H, W = int(ref.shape[0]), int(ref.shape[1])
        _lib.call("d2ic_cg_set_reference", self.handle, _lib.ptr(ref), H, W, _lib.INTERP[interpolation],
                  _lib.stream())
        self._ref_key = key
        self._ref = ref
        return ref

    def close(self):
        """Free the plan's device memory now (also done on garbage collection)."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            _lib.load().d2ic_cg_plan_destroy(h)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        if sys.is_finalizing():  # the HIP runtime may already be torn down
            return
        try:
            self.close()
        except Exception:
            pass


class SolveGraph:
    """A whole batched solve captured once into a HIP graph and replayed per
    frame (no reference counterpart: the reference re-enters its jitted
    ``_cg_solve`` for every frame).

    The graph holds the launch sequence of ``d2ic_cg_solve_batch_noalloc`` --
    descriptor upload, workspace clear, per-problem init, the persistent NLCG
    kernel that runs every iteration, finalize -- for fixed buffers: the
    caller writes the next frame into ``images[p]`` and the warm start into
    ``u[p]`` (in place), then calls :meth:`replay`; ``n_iters``,
    ``pass_count`` and ``hist`` are refreshed by every replay.  Up to 8
    problems (the allocation-free entry point's scratch array).  Every solve
    on a GPU that uses that array (replays of any graph, direct
    ``d2ic_cg_solve_batch_noalloc`` calls) must be serialised, and a replay
    must not run beside other kernels that hold CUs: the captured persistent
    kernel is an ordinary launch whose grid must be co-resident.
    """

    def __init__(self, plans, images, reg_strength=0.1, max_iter: int = 200, tol: float = 1e-3, u=None,
                 with_history: bool = False):
        import ctypes as _ct

        t = _lib.torch()
        plans = list(plans)
        P = len(plans)
        if not 1 <= P <= 8:
            raise ValueError("a solve graph holds 1 to 8 problems")
        if len(images) != P:
            raise ValueError("one image per plan")
        for pl in plans:
            if pl._ref_key is None:
                raise RuntimeError("set_reference() must be called on every plan before capture")
        self.plans = plans
        self.images = [_lib.to_dev(g, t.float32).contiguous().clone() for g in images]
        for g in self.images:
            if g.dim() != 2:
                raise ValueError("images must be 2-D")
        self.u = [t.zeros((pl.n_nodes, 2), dtype=t.float32, device=_lib.device()) for pl in plans]
        if u is not None:
            for dst, src in zip(self.u, u):
                dst.copy_(_lib.to_dev(src, t.float32, dst.shape))
        self.hist = [t.empty((max(int(max_iter), 1), 3), dtype=t.float32, device=_lib.device())
                     if with_history else None for _ in plans]
        self.counters = [t.zeros(2, dtype=t.int32, device=_lib.device()) for _ in plans]
        self.max_iter = int(max_iter)
        vp = _ct.c_void_p
        arr = lambda xs: (vp * P)(*[vp(x) for x in xs])  # noqa: E731
        self._args = (P, arr([pl.handle.value for pl in plans]), arr([g.data_ptr() for g in self.images]),
                      (_ct.c_int32 * P)(*[g.shape[0] for g in self.images]),
                      (_ct.c_int32 * P)(*[g.shape[1] for g in self.images]),
                      arr([x.data_ptr() for x in self.u]),
                      (_ct.c_float * P)(*([float(reg_strength)] * P) if np.isscalar(reg_strength)
                                        else [float(v) for v in reg_strength]),
                      (_ct.c_int32 * P)(*([self.max_iter] * P)), (_ct.c_float * P)(*([float(tol)] * P)),
                      arr([_lib.ptr(h) for h in self.hist]), arr([c.data_ptr() for c in self.counters]),
                      arr([c.data_ptr() + 4 for c in self.counters]))
        # one eager call validates the arguments and resolves the scratch
        # array's device address before capture; the warm start is restored
        u0 = [x.clone() for x in self.u]
        _lib.call("d2ic_cg_solve_batch_noalloc", *self._args, _lib.stream())
        for dst, src in zip(self.u, u0):
            dst.copy_(src)
        t.cuda.synchronize()
        self.graph = t.cuda.CUDAGraph()
        side = t.cuda.Stream()
        side.wait_stream(t.cuda.current_stream())
        with t.cuda.graph(self.graph, stream=side):
            _lib.call("d2ic_cg_solve_batch_noalloc", *self._args, int(side.cuda_stream))
        t.cuda.current_stream().wait_stream(side)

    def replay(self):
        """Run the captured solve on the current contents of ``images`` and ``u``."""
        self.graph.replay()

    def results(self):
        """[(u, n_iters, pass_count)] of the last replay (synchronises)."""
        return [(u, int(c[0]), int(c[1])) for u, c in zip(self.u, (c.cpu() for c in self.counters))]


class GlobalCGSolver(SolverBase):
    """See module docstring.  Same constructor as the reference
    (``solver_global_cg.py:45-63``)."""

    def __init__(self, verbose: bool = False, interpolation: str = "cubic", use_map_coordinates: bool | None = None,
                 log_cg: bool = False) -> None:
        self._compiled = False
        self._plan: CGPlan | None = None
        self._verbose = bool(verbose)
        self._log_cg = bool(log_cg)
        if use_map_coordinates is not None:
            warnings.warn(
                "use_map_coordinates is kept for backward compatibility; use interpolation='cubic' or 'linear'.",
                DeprecationWarning,
                stacklevel=2,
            )
            interpolation = "cubic" if use_map_coordinates else "linear"
        self._interpolation = _normalize_interpolation(interpolation)
        self.last_pass_count: int | None = None

    @property
    def interpolation(self) -> str:
        return self._interpolation

    def compile(self, assets: Any) -> None:
        pixel_data = getattr(assets, "pixel_data", None)
        if pixel_data is None:
            raise ValueError("MeshAssets must provide pixel_data for GlobalCGSolver.")
        self._plan = CGPlan(pixel_data)
        self._compiled = True

    def warmup(self, state: Any) -> None:
        """Samples f at the pixels once; there is nothing to JIT-compile."""
        if not self._compiled or self._plan is None:
            raise RuntimeError("GlobalCGSolver.compile() must be called before warmup().")
        if state.assets.pixel_data is None:
            raise ValueError("MeshAssets must provide pixel_data for GlobalCGSolver warmup.")
        self._plan.set_reference(state.ref_image, self._interpolation)

    @property
    def plan(self) -> CGPlan | None:
        return self._plan

    def autotune_grid(self, state: Any, def_image: Array, **kwargs) -> dict:
        """Pick the persistent grid size for this mesh from short timed solves
        on ``def_image`` (``CGPlan.autotune``); later solves use it.  No
        reference counterpart: launch tuning only, results are unchanged."""
        if not self._compiled or self._plan is None:
            raise RuntimeError("GlobalCGSolver.compile() must be called before autotune_grid().")
        self._plan.set_reference(state.ref_image, self._interpolation)
        kwargs.setdefault("alpha_reg", float(state.config.reg_strength))
        return self._plan.autotune(def_image, **kwargs)

    def set_tiling(self, state: Any, tile_px: int = 256, strip_w: int = 16) -> None:
        """Rebuild the plan with another tiling (``CGPlan``); (256, 16) is the
        default.  The persistent grid returns to the work-based default."""
        if not self._compiled or self._plan is None:
            raise RuntimeError("GlobalCGSolver.compile() must be called before set_tiling().")
        if (int(tile_px), int(strip_w)) != self._plan.tiling:
            self._replace_plan(CGPlan(state.assets.pixel_data, int(tile_px), int(strip_w)), state)

    def autotune_tiling(self, state: Any, def_image: Array, candidates=None, max_iter: int = 8, repeats: int = 3,
                        alpha_reg: float | None = None, cache=None) -> dict:
        """Pick the plan's tiling (``CGPlan`` ``tile_px``, ``strip_w``) for
        this mesh from short timed solves on ``def_image``: each candidate
        plan is built and runs ``repeats`` solves of ``max_iter`` iterations
        (tol 0, from u = 0, default grid); the fastest is kept as the solver's
        plan, the others are freed.  Candidates default to tiles of 256, 192
        and 128 pixels with 16-px strips and 256-px tiles with 12- and 8-px
        strips; the default tiling (256, 16) is always timed.  Tilings differ
        in how tile partials are grouped, so the kept plan's results differ
        from the default plan's by rounding only.  A grid chosen before is
        reset (tune the grid after the tiling: ``autotune``).

        ``cache``: as for the grid autotuner (same file; keyed on device,
        interpolation, image shape, pixels, nodes, candidates, iterations).
        Returns ``{"tiling": (tile_px, strip_w), "ms": {"tile_px/strip_w":
        median ms}, "default_ms": ms of (256, 16), "cached": bool}``.  No
        reference counterpart: launch tuning only."""
        if not self._compiled or self._plan is None:
            raise RuntimeError("GlobalCGSolver.compile() must be called before autotune_tiling().")
        t = _lib.torch()
        pix = state.assets.pixel_data
        if alpha_reg is None:
            alpha_reg = float(state.config.reg_strength)
        default = CGPlan.DEFAULT_TILING
        if candidates is None:
            candidates = [(256, 16), (192, 16), (128, 16), (256, 12), (256, 8)]
        cands = [default] + [tuple(int(v) for v in c) for c in candidates if tuple(int(v) for v in c) != default]
        name = lambda c: f"{c[0]}/{c[1]}"  # noqa: E731
        ref = self._plan.set_reference(state.ref_image, self._interpolation)
        g = _lib.to_dev(def_image, t.float32)
        path = _autotune_cache_path(cache)
        key = None
        if path is not None:
            key = "|".join(str(v) for v in (
                "tiling", t.cuda.get_device_name(_lib.device()), self._interpolation,
                "x".join(map(str, ref.shape)), self._plan.n_pixels, self._plan.n_nodes,
                ",".join(map(name, cands)), int(max_iter)))
            hit = _autotune_cache_load(path).get(key)
            if hit is not None and tuple(hit["tiling"]) in cands:
                best = tuple(int(v) for v in hit["tiling"])
                if best != self._plan.tiling:
                    self._replace_plan(CGPlan(pix, *best), state)
                ms = {str(k): float(v) for k, v in hit["ms"].items()}
                return {"tiling": best, "ms": ms, "default_ms": ms.get(name(default)), "cached": True}
        plans = {self._plan.tiling: self._plan}
        times = {}
        for c in cands:
            if c not in plans:
                plans[c] = CGPlan(pix, *c)
                plans[c].set_reference(state.ref_image, self._interpolation)
            plans[c].set_grid(0)
            times[c] = plans[c].time_solves(g, alpha_reg, max_iter, repeats)
        best = min(times, key=times.get)
        keep = plans.pop(best)
        for p in plans.values():
            if p is not self._plan:
                p.close()
        if keep is not self._plan:
            self._replace_plan(keep, state)
        else:
            keep.set_grid(0)
        ms = {name(c): v for c, v in times.items()}
        if path is not None:
            _autotune_cache_store(path, key, {"tiling": list(best), "ms": ms})
        return {"tiling": best, "ms": ms, "default_ms": times[default], "cached": False}

    def autotune(self, state: Any, def_image: Array, tiling_candidates=None, grid_candidates=None,
                 **kwargs) -> dict:
        """Tiling, then the persistent grid on the kept plan
        (``autotune_tiling``, ``autotune_grid``); the other keyword arguments
        (max_iter, repeats, alpha_reg, cache) go to both."""
        tiling = self.autotune_tiling(state, def_image, candidates=tiling_candidates, **kwargs)
        grid = self.autotune_grid(state, def_image, candidates=grid_candidates, **kwargs)
        return {"tiling": tiling, "grid": grid}

    def _replace_plan(self, plan: CGPlan, state: Any) -> None:
        old = self._plan
        self._plan = plan
        plan.set_reference(state.ref_image, self._interpolation)
        if old is not None and old is not plan:
            old.close()

    def close(self) -> None:
        """Release the device plan."""
        if self._plan is not None:
            self._plan.close()
            self._plan = None
            self._compiled = False

    def solve(self, state: Any, def_image: Array) -> GlobalCGResult:
        res = self.launch(state, def_image).finish()
        return dataclasses.replace(res, strain=_lib.torch().zeros_like(res.u_nodal))  # the reference's placeholder

    def launch(self, state: Any, def_image: Array) -> "PendingSolve":
        """Queue the solve on the current stream without waiting for it: the
        returned PendingSolve holds the result arrays (u is filled in place by
        the queued kernels; no placeholder strain) and finish() reads the
        iteration count (the only host wait of a solve)."""
        if not self._compiled or self._plan is None:
            raise RuntimeError("GlobalCGSolver.compile() must be called before solve().")
        t = _lib.torch()
        plan = self._plan
        plan.set_reference(state.ref_image, self._interpolation)
        g = _lib.to_dev(def_image, t.float32)
        if g.dim() != 2:
            raise ValueError("def_image must be a 2-D array (H, W)")
        nn = plan.n_nodes
        if state.u0_nodal is not None:
            u = _lib.to_dev(state.u0_nodal, t.float32, (nn, 2)).clone()
        else:
            u = t.zeros((nn, 2), dtype=t.float32, device=_lib.device())
        cfg = state.config
        max_iter = int(cfg.max_iters)
        hist = t.empty((max(max_iter, 0), 3), dtype=t.float32, device=_lib.device())
        counters = t.zeros((2,), dtype=t.int32, device=_lib.device())
        _lib.call("d2ic_cg_solve", plan.handle, _lib.ptr(g), int(g.shape[0]), int(g.shape[1]), _lib.ptr(u),
                  float(cfg.reg_strength), max_iter, float(cfg.tol), _lib.ptr(hist), _lib.ptr(counters),
                  _lib.ptr(counters) + 4, _lib.stream())
        save_history = bool(getattr(cfg, "save_history", False))
        result = GlobalCGResult(u_nodal=u, strain=None, history=hist if save_history else None, n_iters=None)
        return PendingSolve(self, result, hist, counters, g)

    def _finish(self, pending: "PendingSolve") -> GlobalCGResult:
        n_iters, passes = (int(v) for v in pending.counters.cpu().tolist())
        if n_iters < 0:
            raise RuntimeError("d2ic_cg_solve did not finish (a grid barrier timed out)")
        self.last_pass_count = passes
        if self._verbose and self._log_cg:
            h = pending.hist[:max(n_iters + 1, 1)].cpu().numpy()
            for k, (J, gn, a) in enumerate(h):
                if np.isfinite(J):
                    print(f"CG iter {k}: J={J:.3e}, |grad|={gn:.3e}, alpha={a:.3e}")
        pending.result = dataclasses.replace(pending.result, n_iters=n_iters)
        return pending.result
