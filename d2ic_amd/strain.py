"""Nodal Green-Lagrange strain by weighted least squares on the GPU.

Reference ``d2ic/strain.py:11-148``: for each node, A = sum w dX dX^T + eps I,
grad u_x = A^-1 sum w dX du_x (likewise u_y), F = I + grad u,
E = (F^T F - I)/2, Voigt [E11, E22, (E12 + E21)/2]; w = 1 or
exp(-(|dX|/L)^2) when gauge_length L > 0.  Kernel: ``d2ic_strain``.
"""
from __future__ import annotations

from typing import Tuple

from . import _lib
from .types import Array


def _strain(displacement, nodes_coord, node_neighbor_index, node_neighbor_degree, gauge_length, eps):
    t = _lib.require_gpu()
    u = _lib.to_dev(displacement, t.float32).reshape(-1, 2)
    X = _lib.to_dev(nodes_coord, t.float32).reshape(-1, 2)
    idx = _lib.to_dev(node_neighbor_index, t.int32)
    deg = _lib.to_dev(node_neighbor_degree, t.int32)
    n = int(X.shape[0])
    md = int(idx.shape[1]) if idx.dim() == 2 else 0
    F = t.empty((n, 2, 2), dtype=t.float32, device=_lib.device())
    E = t.empty((n, 2, 2), dtype=t.float32, device=_lib.device())
    V = t.empty((n, 3), dtype=t.float32, device=_lib.device())
    _lib.call("d2ic_strain", _lib.ptr(u), _lib.ptr(X), n, _lib.ptr(idx), _lib.ptr(deg), md, float(gauge_length),
              float(eps), _lib.ptr(F), _lib.ptr(E), _lib.ptr(V), _lib.stream())
    return F, E, V


def compute_green_lagrange_strain_nodes_lsq(displacement: Array, nodes_coord: Array, node_neighbor_index: Array,
                                            node_neighbor_degree: Array, gauge_length: float = 0.0,
                                            eps: float = 1e-8) -> Tuple[Array, Array]:
    """Returns (F_all, E_all), each (Nn, 2, 2) (``strain.py:11-92``)."""
    F, E, _ = _strain(displacement, nodes_coord, node_neighbor_index, node_neighbor_degree, gauge_length, eps)
    return F, E


def green_lagrange_to_voigt(E_all: Array) -> Array:
    """(N,2,2) -> (N,3) [E11, E22, E12] with E12 = (E01 + E10)/2 (``strain.py:95-114``)."""
    t = _lib.torch()
    E = E_all if isinstance(E_all, t.Tensor) else t.as_tensor(E_all)
    sym = 0.5 * (E + E.transpose(1, 2))
    return t.stack([sym[:, 0, 0], sym[:, 1, 1], sym[:, 0, 1]], dim=1)


def small_strain_nodes_lsq(displacement: Array, nodes_coord: Array, node_neighbor_index: Array,
                           node_neighbor_degree: Array, gauge_length: float = 0.0, eps: float = 1e-8) -> Array:
    """Green-Lagrange strain in Voigt form (``strain.py:117-148``)."""
    _, _, V = _strain(displacement, nodes_coord, node_neighbor_index, node_neighbor_degree, gauge_length, eps)
    return V


def strain_voigt(displacement, nodes_coord, node_neighbor_index, node_neighbor_degree, gauge_length=0.0,
                 eps=1e-8) -> Array:
    """Single-kernel form of compute + green_lagrange_to_voigt."""
    return _strain(displacement, nodes_coord, node_neighbor_index, node_neighbor_degree, gauge_length, eps)[2]
