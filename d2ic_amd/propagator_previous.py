"""u_warm = u_prev (reference ``d2ic/propagator_previous.py:9-20``)."""
from __future__ import annotations

from typing import Optional

from .propagator_base import DisplacementPropagatorBase
from .types import Array


class PreviousDisplacementPropagator(DisplacementPropagatorBase):
    def propagate(self, u_prev: Optional[Array], u_prevprev: Optional[Array] = None) -> Optional[Array]:
        return u_prev
