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


def test_grid_autotune_cache_file(tmp_path, monkeypatch):
    """The persisted grid-autotune cache: default path from the environment,
    disabled with False, entries added without losing earlier ones, and a
    corrupt file read as empty."""
    from d2ic_amd import solver_global_cg as m

    monkeypatch.setenv("D2IC_AUTOTUNE_CACHE", str(tmp_path / "c.json"))
    assert m._autotune_cache_path(None) == str(tmp_path / "c.json")
    assert m._autotune_cache_path(False) is None
    p = str(tmp_path / "sub" / "cache.json")
    assert m._autotune_cache_load(p) == {}
    m._autotune_cache_store(p, "a", {"blocks": 64, "ms": {"0": 1.0, "64": 0.5}})
    m._autotune_cache_store(p, "b", {"blocks": 0, "ms": {"0": 0.2}})
    d = m._autotune_cache_load(p)
    assert set(d) == {"a", "b"} and d["a"]["blocks"] == 64
    (tmp_path / "bad.json").write_text("{not json")
    assert m._autotune_cache_load(str(tmp_path / "bad.json")) == {}
