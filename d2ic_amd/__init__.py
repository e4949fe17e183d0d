"""d2ic_amd -- MI355X-native global FE-DIC hot path with the D2IC API.

Import root for the drop-in classes of the reference package ``d2ic``
(``d2ic/__init__.py:1-81``) that this package covers: the global NLCG solver,
ZNCC initialisation, the mesh pipeline, the sequence runner, propagators,
pixel assets, the structured mesher and strain.  Compute runs in the HIP
library ``libd2ic_amd.so`` (see ``d2ic_amd/csrc``).
"""
from .types import Array
from .dataclasses import (
    InitMotionConfig,
    MeshDICConfig,
    DICResult,
    DICDiagnostics,
    BatchConfig,
    BatchResult,
    BatchDiagnostics,
)
from .mesh_assets import Mesh, MeshAssets, PixelAssets, compute_element_centers, make_mesh_assets, \
    build_node_neighbor_tables
from .mask2mesh import downsample_mask, mask_to_mesh, mask_to_mesh_assets, mask_to_mesh_gmsh, \
    mask_to_mesh_assets_gmsh
from .pixel_assets import build_pixel_assets
from .strain import compute_green_lagrange_strain_nodes_lsq, green_lagrange_to_voigt, small_strain_nodes_lsq
from .dic_init_motion import DICInitMotion
from .dic_mesh_based import DICMeshBased
from .batch_base import BatchBase
from .batch_mesh_based import BatchMeshBased
from .propagator_base import DisplacementPropagatorBase
from .propagator_previous import PreviousDisplacementPropagator
from .propagator_constant_velocity import ConstantVelocityPropagator
from .solver_translation_zncc import TranslationZNCCSolver
from .solver_global_cg import GlobalCGSolver
from .solver_local_gn import LocalGaussNewtonSolver, LocalGNResult

__all__ = [
    "Array",
    "InitMotionConfig",
    "MeshDICConfig",
    "DICResult",
    "DICDiagnostics",
    "BatchConfig",
    "BatchResult",
    "BatchDiagnostics",
    "Mesh",
    "MeshAssets",
    "PixelAssets",
    "compute_element_centers",
    "make_mesh_assets",
    "build_node_neighbor_tables",
    "downsample_mask",
    "mask_to_mesh",
    "mask_to_mesh_assets",
    "mask_to_mesh_gmsh",
    "mask_to_mesh_assets_gmsh",
    "build_pixel_assets",
    "compute_green_lagrange_strain_nodes_lsq",
    "green_lagrange_to_voigt",
    "small_strain_nodes_lsq",
    "DICInitMotion",
    "DICMeshBased",
    "BatchBase",
    "BatchMeshBased",
    "DisplacementPropagatorBase",
    "PreviousDisplacementPropagator",
    "ConstantVelocityPropagator",
    "TranslationZNCCSolver",
    "GlobalCGSolver",
    "LocalGaussNewtonSolver",
    "LocalGNResult",
]
