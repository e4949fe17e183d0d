"""Constructor-level argument checks of the drop-in classes (no GPU needed):
the reference's ValueErrors for unsupported interpolation
(solver_global_cg.py:472-478) and even ZNCC windows
(solver_translation_zncc.py:39-40), and this package's layout option."""
from __future__ import annotations

import pytest


def test_constructor_errors():
    from d2ic_amd import GlobalCGSolver, LocalGaussNewtonSolver
    from d2ic_amd.dataclasses import InitMotionConfig
    from d2ic_amd.solver_translation_zncc import TranslationZNCCSolver

    with pytest.raises(ValueError):
        GlobalCGSolver(interpolation="quintic")
    with pytest.raises(ValueError):
        LocalGaussNewtonSolver(interpolation="nearest")
    with pytest.raises(ValueError):
        LocalGaussNewtonSolver(layout="rows")
    with pytest.raises(ValueError):
        TranslationZNCCSolver(InitMotionConfig(win=20))
    for name in ("linear", "bilinear", "cubic"):
        GlobalCGSolver(interpolation=name)
        LocalGaussNewtonSolver(interpolation=name)
