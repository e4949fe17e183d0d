"""Mesh-based DIC pipeline (reference ``d2ic/dic_mesh_based.py:30-155``).

``prepare`` builds missing neighbour tables and the pixel assets (on the GPU),
puts the reference image and nodes on the device and prepares the solver;
``run`` solves one deformed image, then computes nodal Green-Lagrange strain
and, if configured, the discrepancy map.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Optional

from . import _lib
from .dataclasses import DICDiagnostics, DICResult, MeshDICConfig
from .dic_base import DICBase
from .discrepancy import compute_discrepancy_map_ref
from .mesh_assets import Mesh, MeshAssets, build_node_neighbor_tables, make_mesh_assets
from .pixel_assets import build_pixel_assets
from .solver_base import SolverBase
from .strain import strain_voigt
from .types import Array


@dataclass
class MeshDICState:
    """``dic_mesh_based.py:30-38``."""
    ref_image: Array
    nodes_xy_device: Array
    assets: MeshAssets
    config: MeshDICConfig
    u0_nodal: Array | None = None


class DICMeshBased(DICBase):
    def __init__(self, mesh: Mesh, solver: SolverBase, config: MeshDICConfig) -> None:
        self.mesh = mesh
        self.solver = solver
        self.config = config
        self._state: MeshDICState | None = None
        self._nbr_dev = None  # device copies of the strain neighbour tables (made in prepare)

    @property
    def state(self) -> MeshDICState | None:
        return self._state

    def prepare(self, ref_image: Array, assets: Optional[MeshAssets] = None) -> None:
        t = _lib.require_gpu()
        if assets is None:
            assets = make_mesh_assets(self.mesh, with_neighbors=True)
        elif assets.node_neighbor_index is None or assets.node_neighbor_degree is None:
            idx, deg = build_node_neighbor_tables(assets.mesh)
            assets = replace(assets, node_neighbor_index=idx, node_neighbor_degree=deg)
        if assets.pixel_data is None:
            pixel_data = build_pixel_assets(mesh=assets.mesh, ref_image=ref_image, binning=1.0,
                                            roi_mask=getattr(assets, "roi_mask", None))
            assets = replace(assets, pixel_data=pixel_data)
        ref_dev = _lib.to_dev(ref_image, t.float32)
        nodes_dev = _lib.to_dev(assets.mesh.nodes_xy, t.float32)
        self._state = MeshDICState(ref_image=ref_dev, nodes_xy_device=nodes_dev, assets=assets,
                                   config=self.config, u0_nodal=None)
        self._nbr_dev = None
        if assets.node_neighbor_index is not None and assets.node_neighbor_degree is not None:
            # one host->device copy per prepare instead of one per frame
            self._nbr_dev = (_lib.to_dev(assets.node_neighbor_index, t.int32),
                             _lib.to_dev(assets.node_neighbor_degree, t.int32))
        self.solver.compile(assets)
        if hasattr(self.solver, "warmup"):
            self.solver.warmup(self._state)

    def set_initial_guess(self, u0_nodal: Array) -> None:
        if self._state is None:
            raise RuntimeError("Call prepare() before set_initial_guess().")
        self._state.u0_nodal = u0_nodal

    def run(self, def_image: Array) -> DICResult:
        if self._state is None:
            raise RuntimeError("DICMeshBased.prepare() must be called before run().")
        t = _lib.torch()
        # the HIP solver's launches are queued without waiting for the solve,
        # so the strain kernel is queued behind them before the host reads
        # the iteration count
        launch = getattr(self.solver, "launch", None)
        pending = launch(self._state, def_image) if launch is not None else None
        sol = pending.result if pending is not None else self.solver.solve(self._state, def_image)
        u_nodal = sol.u_nodal
        assets = self._state.assets
        info = {"stage": "mesh_based"}
        pixel_maps = {}
        strain = strain_info = strain_error = None
        if assets.node_neighbor_index is not None and assets.node_neighbor_degree is not None:
            idx, deg = self._nbr_dev if self._nbr_dev is not None else (assets.node_neighbor_index,
                                                                       assets.node_neighbor_degree)
            try:
                strain = strain_voigt(u_nodal, self._state.nodes_xy_device, idx, deg,
                                      gauge_length=self.config.strain_gauge_length, eps=self.config.strain_eps)
                strain_info = "green_lagrange"
            except Exception as exc:  # pragma: no cover - defensive, as in the reference
                strain = t.zeros((u_nodal.shape[0], 3), dtype=u_nodal.dtype, device=u_nodal.device)
                strain_info = "failed_green_lagrange"
                strain_error = str(exc)
        if pending is not None:
            sol = pending.finish()
        if getattr(sol, "n_iters", None) is not None:
            info["n_iters"] = int(sol.n_iters)
        if getattr(sol, "history", None) is not None:
            info["history"] = "attached"
        if strain_info is not None:
            info["strain"] = strain_info
            if strain_error is not None:
                info["strain_error"] = strain_error
        else:
            strain = t.zeros((u_nodal.shape[0], 3), dtype=u_nodal.dtype, device=u_nodal.device)
            info["strain"] = "skipped_no_neighbors"
        if bool(getattr(self.config, "compute_discrepancy_map", False)):
            if assets.pixel_data is None:
                info["discrepancy"] = "skipped_no_pixel_assets"
            else:
                disc, rms = compute_discrepancy_map_ref(ref_image=self._state.ref_image, def_image=def_image,
                                                        u_nodal=u_nodal, pixel_assets=assets.pixel_data)
                pixel_maps["discrepancy_ref"] = disc
                info["discrepancy"] = "discrepancy_ref"
                info["discrepancy_rms"] = float(rms)
        return DICResult(u_nodal=u_nodal, strain=strain, diagnostics=DICDiagnostics(info=info),
                         history=getattr(sol, "history", None), pixel_maps=pixel_maps)
