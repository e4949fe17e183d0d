"""Configuration and result dataclasses.

Same field names, defaults and semantics as the reference's
``d2ic/dataclasses.py:8-239`` so that callers only change the import root.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, Optional, Sequence

from .types import Array


@dataclass(frozen=True)
class InitMotionConfig:
    """Element-centre ZNCC initialisation (``dataclasses.py:8-35``).

    win: odd square template size; search: max integer shift per axis;
    score_min: reject matches below this ZNCC score (None disables);
    max_centers: optional cap on the number of centres evaluated.
    """
    win: int = 41
    search: int = 24
    score_min: Optional[float] = 0.2
    max_centers: Optional[int] = None


@dataclass(frozen=True)
class MeshDICConfig:
    """Global mesh DIC refinement (``dataclasses.py:38-72``).

    max_iters / tol: NLCG iteration cap and stop threshold on the 2-norm of
    the objective gradient; reg_strength: spring regularisation weight;
    strain_gauge_length / strain_eps: strain LSQ weighting and Tikhonov term;
    save_history: keep the per-iteration [J, |grad|, alpha] history;
    compute_discrepancy_map: also return the reference-domain residual map.
    """
    max_iters: int = 50
    tol: float = 1e-6
    reg_strength: float = 0.0
    strain_gauge_length: float = 0.0
    strain_eps: float = 1e-8
    save_history: bool = True
    compute_discrepancy_map: bool = False


@dataclass(frozen=True)
class DICDiagnostics:
    """Free-form diagnostics (``dataclasses.py:75-90``)."""
    info: Dict[str, Any] = field(default_factory=dict)


@dataclass(frozen=True)
class DICResult:
    """Standard DIC output (``dataclasses.py:93-121``): nodal u (Nn,2),
    Voigt strain (Nn,3), diagnostics, optional history and user maps."""
    u_nodal: Array
    strain: Array
    diagnostics: DICDiagnostics
    history: Array | None = None
    fields: Dict[str, Array] = field(default_factory=dict)
    pixel_maps: Dict[str, Array] = field(default_factory=dict)


@dataclass(frozen=True)
class BatchConfig:
    """Sequence processing options (``dataclasses.py:129-209``).  The
    ``use_init_motion`` / ``init_motion_*`` flags are declared but, as in the
    reference, not read by ``BatchMeshBased``."""
    use_init_motion: bool = True
    warm_start_from_previous: bool = True
    init_motion_first_frame_only: bool = False
    init_motion_every_frame: bool = False
    prefer_init_motion_over_propagation: bool = False
    verbose: bool = False
    progress: bool = False
    save_per_frame: bool = False
    per_frame_dir: Optional[str] = None
    export_png: bool = False
    export_frames: Sequence[int] | None = None
    png_dir: Optional[str] = None
    plot_fields: Sequence[str] = ("u1", "u2", "e11", "e22", "e12")
    plot_include_discrepancy: bool = False
    plot_cmap: str = "jet"
    plot_alpha: float = 0.6
    plot_mesh: bool = True
    plot_dpi: int = 200
    plot_binning: float = 1.0
    plot_projection: str | bool = "fast"


@dataclass(frozen=True)
class BatchDiagnostics:
    """Batch-level diagnostics (``dataclasses.py:212-223``)."""
    info: Dict[str, Any] = field(default_factory=dict)


@dataclass(frozen=True)
class BatchResult:
    """Per-frame results plus diagnostics (``dataclasses.py:226-239``)."""
    results: list
    diagnostics: BatchDiagnostics
