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


def test_pending_solve_finish_fills_iterations_once():
    """GlobalCGSolver.launch / PendingSolve.finish (CPU stand-in for the device
    counters): finish reads the two counters once, records the pass count,
    returns a result carrying n_iters, and raises when the solve reported a
    barrier timeout (n_iters < 0)."""
    import numpy as np
    import pytest

    from d2ic_amd.solver_global_cg import GlobalCGResult, GlobalCGSolver, PendingSolve

    class Counters:
        def __init__(self, vals):
            self.vals = vals
            self.reads = 0

        def cpu(self):
            self.reads += 1
            return self

        def tolist(self):
            return list(self.vals)

    solver = GlobalCGSolver.__new__(GlobalCGSolver)
    solver._verbose = False
    solver._log_cg = False
    solver.last_pass_count = None
    res = GlobalCGResult(u_nodal=np.zeros((4, 2), np.float32), strain=None, history=None, n_iters=None)
    c = Counters([17, 18])
    p = PendingSolve(solver, res, None, c, object())
    out = p.finish()
    assert out.n_iters == 17 and solver.last_pass_count == 18
    assert p.finish() is out and c.reads == 1
    bad = PendingSolve(solver, res, None, Counters([-1, 3]), None)
    with pytest.raises(RuntimeError, match="did not finish"):
        bad.finish()
