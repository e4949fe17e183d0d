"""Sequential frame runner (reference ``d2ic/batch_mesh_based.py:25-240``).

Per frame: warm start from a propagator (or the previous solution when
``warm_start_from_previous``), solve through the injected ``DICMeshBased``,
optionally chain a local pipeline, optionally save ``frame_{k:04d}.npz`` with
keys ``u_nodal``, ``strain`` (and ``discrepancy_ref``).  Warm-start fields stay
on the GPU.  PNG export (matplotlib plotting) is outside this package.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from .batch_base import BatchBase
from .dataclasses import BatchConfig, BatchDiagnostics, BatchResult, DICResult
from .dic_mesh_based import DICMeshBased
from .mesh_assets import MeshAssets
from .propagator_base import DisplacementPropagatorBase
from .types import Array


@dataclass
class BatchState:
    ref_image: Array
    assets: MeshAssets
    is_prepared: bool = False


class BatchMeshBased(BatchBase):
    def __init__(self, ref_image: Array, assets: MeshAssets, dic_mesh: DICMeshBased, batch_config: BatchConfig,
                 dic_local: Optional[DICMeshBased] = None,
                 propagator: Optional[DisplacementPropagatorBase] = None) -> None:
        super().__init__()
        self.ref_image = ref_image
        self.assets = assets
        self.dic_mesh = dic_mesh
        self.dic_local = dic_local
        self.config = batch_config
        self.propagator = propagator
        self._state = BatchState(ref_image=ref_image, assets=assets, is_prepared=False)

    def before(self, images: Sequence[Array]) -> None:
        self.dic_mesh.prepare(self.ref_image, self.assets)
        if self.dic_local is not None:
            self.dic_local.prepare(self.ref_image, self.assets)
        self._state.is_prepared = True

    def sequence(self, images: Sequence[Array] | Iterable[Array]) -> BatchResult:
        if not self._state.is_prepared:
            raise RuntimeError("BatchMeshBased.before() must be called before sequence().")
        if bool(getattr(self.config, "export_png", False)):
            raise NotImplementedError("PNG export (plotting) is not part of d2ic_amd.")
        t = _lib.torch()
        save = bool(getattr(self.config, "save_per_frame", False))
        out_dir = None
        if save:
            out_dir = Path(self.config.per_frame_dir) if self.config.per_frame_dir else (
                Path.cwd() / "_outputs" / "per_frame_fields")
            out_dir.mkdir(parents=True, exist_ok=True)
        progress = bool(getattr(self.config, "progress", False))
        verbose = bool(getattr(self.config, "verbose", False))
        try:
            n_frames = len(images)  # type: ignore[arg-type]
        except TypeError:
            n_frames = None
        per_frame: list[DICResult] = []
        u_prev = u_prevprev = None
        for k, Idef in enumerate(images):
            if progress or verbose:
                print(f"[Batch] Frame {k + 1}{'' if n_frames is None else '/' + str(n_frames)}: start")
            if self.propagator is not None:
                u_warm = self.propagator.propagate(u_prev=u_prev, u_prevprev=u_prevprev)
            else:
                u_warm = u_prev if self.config.warm_start_from_previous else None
            if u_warm is not None:
                self.dic_mesh.set_initial_guess(_lib.to_dev(u_warm, t.float32).clone())
            res = self.dic_mesh.run(Idef)
            if self.dic_local is not None:
                self.dic_local.set_initial_guess(res.u_nodal.clone())
                res = self.dic_local.run(Idef)
            per_frame.append(res)
            if save:
                payload = {"u_nodal": _lib.to_np(res.u_nodal), "strain": _lib.to_np(res.strain)}
                disc = res.pixel_maps.get("discrepancy_ref") if isinstance(res.pixel_maps, dict) else None
                if disc is not None:
                    payload["discrepancy_ref"] = _lib.to_np(disc)
                np.savez_compressed(out_dir / f"frame_{k:04d}.npz", **payload)
            u_prevprev, u_prev = u_prev, res.u_nodal
            if progress or verbose:
                print(f"[Batch] Frame {k + 1}{'' if n_frames is None else '/' + str(n_frames)}: done")
        diag = BatchDiagnostics(info={"stage": "batch_mesh_based", "n_frames": len(per_frame),
                                      "warm_start_from_previous": self.config.warm_start_from_previous})
        return BatchResult(results=per_frame, diagnostics=diag)

    def end(self, result: BatchResult) -> BatchResult:
        return BatchResult(results=result.results,
                           diagnostics=BatchDiagnostics(info={**result.diagnostics.info, "post": "end() placeholder"}))
