"""Translation-only ZNCC matching at element centres on the GPU.

Drop-in for ``TranslationZNCCSolver`` (reference
``d2ic/solver_translation_zncc.py:26-168``): for every centre, the (2s+1)^2
integer shifts of a win x win template are scored by zero-mean normalised
cross-correlation and the first maximum is kept.  One HIP workgroup per
centre stages the normalised template and the search patch in LDS
(``d2ic_zncc_match``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from . import _lib
from .dataclasses import InitMotionConfig
from .solver_base import SolverBase
from .types import Array


@dataclass(frozen=True)
class TranslationZNCCResult:
    """``solver_translation_zncc.py:17-23``."""
    u_centers: Array  # (Ne, 2)
    scores: Array  # (Ne,)


def match_centers(centers, ref_image, def_image, *, win: int, search: int):
    """``_match_centers_jax`` (``solver_translation_zncc.py:118-168``) on device tensors."""
    t = _lib.require_gpu()
    if int(win) % 2 == 0 or int(win) <= 0:
        raise ValueError("win must be odd and positive")
    c = _lib.to_dev(centers, t.float32).reshape(-1, 2)
    r = _lib.to_dev(ref_image, t.float32)
    d = _lib.to_dev(def_image, t.float32)
    n = int(c.shape[0])
    u = t.empty((n, 2), dtype=t.float32, device=_lib.device())
    s = t.empty((n,), dtype=t.float32, device=_lib.device())
    _lib.call("d2ic_zncc_match", _lib.ptr(c), n, _lib.ptr(r), int(r.shape[0]), int(r.shape[1]), _lib.ptr(d),
              int(d.shape[0]), int(d.shape[1]), int(win), int(search), _lib.ptr(u), _lib.ptr(s), _lib.stream())
    return u, s


class TranslationZNCCSolver(SolverBase):
    """``solver_translation_zncc.py:26-82``; ``chunk_size`` is kept for API parity."""

    def __init__(self, config: InitMotionConfig, chunk_size: int = 128) -> None:
        if config.win % 2 == 0:
            raise ValueError("InitMotionConfig.win must be odd for ZNCC matching.")
        self.config = config
        self.chunk_size = max(1, int(chunk_size))
        self._compiled = False

    def compile(self, assets: Any) -> None:
        self._compiled = True

    def solve(self, state: Any, def_image: Array) -> TranslationZNCCResult:
        if not self._compiled:
            raise RuntimeError("TranslationZNCCSolver.compile() must be called before solve().")
        t = _lib.require_gpu()
        centers = _lib.to_dev(state.assets.element_centers_xy, t.float32).reshape(-1, 2)
        if self.config.max_centers is not None:
            centers = centers[: int(self.config.max_centers)].contiguous()
        u, scores = match_centers(centers, state.ref_image, def_image, win=int(self.config.win),
                                  search=int(self.config.search))
        if self.config.score_min is not None:
            valid = scores >= float(self.config.score_min)
            u = t.where(valid[:, None], u, t.zeros_like(u))
            scores = t.where(valid, scores, t.full_like(scores, float("nan")))
        return TranslationZNCCResult(u_centers=u, scores=scores)
