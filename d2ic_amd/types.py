"""Array type used across the package.

Device-resident arrays are ``torch.Tensor`` objects on the GPU (torch is used
only for device memory and streams); host-side geometry is ``numpy.ndarray``.
Mirrors ``d2ic/types.py:1-13`` of the reference.
"""
from __future__ import annotations

from typing import Any, Tuple, Union

import numpy as np

try:
    import torch

    Array = Union["torch.Tensor", np.ndarray]
except Exception:  # pragma: no cover
    Array = Any

NPArray = np.ndarray
XY = Tuple[float, float]
