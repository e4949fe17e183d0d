"""Coarse initialisation pipeline (reference ``d2ic/dic_init_motion.py:23-91``):
ZNCC shifts at element centres, projected to nodes by global inverse
distance weighting (power 2) on the GPU (``d2ic_idw``)."""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib
from .dataclasses import DICDiagnostics, DICResult, InitMotionConfig
from .dic_base import DICBase
from .mesh_assets import MeshAssets
from .solver_base import SolverBase
from .types import Array


@dataclass
class InitMotionState:
    ref_image: Array
    assets: MeshAssets


def project_centers_to_nodes_idw(nodes_xy, centers_xy, disp_centers):
    """``_project_centers_to_nodes_idw`` (``dic_init_motion.py:64-91``): centres
    whose shift holds a NaN are dropped; a centre within 1e-6 of a node is copied."""
    t = _lib.require_gpu()
    nodes = _lib.to_dev(nodes_xy, t.float32).reshape(-1, 2)
    c = _lib.to_dev(centers_xy, t.float32).reshape(-1, 2)
    d = _lib.to_dev(disp_centers, t.float32).reshape(-1, 2)
    keep = ~t.isnan(d).any(dim=1)
    c = c[keep].contiguous()
    d = d[keep].contiguous()
    out = t.zeros_like(nodes)
    _lib.call("d2ic_idw", _lib.ptr(nodes), int(nodes.shape[0]), _lib.ptr(c), _lib.ptr(d), int(c.shape[0]),
              _lib.ptr(out), _lib.stream())
    return out


class DICInitMotion(DICBase):
    def __init__(self, config: InitMotionConfig, solver: SolverBase) -> None:
        self.config = config
        self.solver = solver
        self._state: InitMotionState | None = None

    def prepare(self, ref_image: Array, assets: MeshAssets) -> None:
        self._state = InitMotionState(ref_image=ref_image, assets=assets)
        self.solver.compile(assets)

    def run(self, def_image: Array) -> DICResult:
        if self._state is None:
            raise RuntimeError("DICInitMotion.prepare() must be called before run().")
        t = _lib.require_gpu()
        sol = self.solver.solve(self._state, def_image)
        centers = self._state.assets.element_centers_xy
        if getattr(self.solver, "config", None) is not None and self.solver.config.max_centers is not None:
            centers = _lib.to_dev(centers, t.float32)[: int(self.solver.config.max_centers)]
        u_nodal = project_centers_to_nodes_idw(self._state.assets.mesh.nodes_xy, centers, sol.u_centers)
        diag = DICDiagnostics(info={"stage": "init_motion", "matches": int((~t.isnan(sol.scores)).sum().item()),
                                    "note": "IDW projection"})
        return DICResult(u_nodal=u_nodal, strain=t.zeros_like(u_nodal), diagnostics=diag, history=None)
