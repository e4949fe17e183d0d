"""u_warm = 2 u_prev - u_prevprev (reference ``d2ic/propagator_constant_velocity.py:10-32``);
falls back to u_prev when only one frame is known and to None before that.
Works on device tensors without leaving the GPU."""
from __future__ import annotations

from typing import Optional

from .propagator_base import DisplacementPropagatorBase
from .types import Array


class ConstantVelocityPropagator(DisplacementPropagatorBase):
    def propagate(self, u_prev: Optional[Array], u_prevprev: Optional[Array] = None) -> Optional[Array]:
        if u_prev is None:
            return None
        if u_prevprev is None:
            return u_prev
        return 2.0 * u_prev - u_prevprev
