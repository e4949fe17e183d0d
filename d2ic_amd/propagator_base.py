"""Warm-start strategy interface (reference ``d2ic/propagator_base.py:9-34``)."""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Optional

from .types import Array


class DisplacementPropagatorBase(ABC):
    @abstractmethod
    def propagate(self, u_prev: Optional[Array], u_prevprev: Optional[Array] = None) -> Optional[Array]:
        """Warm-start displacement (Nn, 2) from the previous frames, or None."""
        raise NotImplementedError
