"""Pins the C oracle's NLCG loop (``orc_cg_solve``) past iteration 0.

A second restatement of ``_cg_solve`` (``d2ic/solver_global_cg.py:326-469``)
in numpy, ``oracle/nlcg.py``, is driven by the C oracle's objective and
gradient (themselves pinned by closed forms and central differences in
test_oracle_objective) and must reproduce the C loop's iterates, n_iters and
[J, |grad|, alpha] history in FP64.  The cases are chosen so that the branches
of the loop all occur: a negative Polak-Ribiere ratio clamped to 0, restarts
on g.d >= 0, Armijo backtracking to depths 2-8, the 2^-10 step applied when
all 10 trials fail, and the stop test before the update.  Planting a wrong rule
in the numpy restatement (Fletcher-Reeves beta, no clamp, no restart, another
fall-through step) must break the agreement, which shows the comparison can
catch such a mistake in the C loop."""
from __future__ import annotations

import functools

import numpy as np
import pytest

from oracle import oracle as O
from oracle.nlcg import beta_prplus, nlcg
from tests._cases import structured_case

# (H, ROI side, element px, speckle seed, ux, uy, reg_strength, u0 noise, interp, max_iter)
CASES = {
    "cubic_rigid": (64, 48, 16, 1, 0.3, -0.2, 0.1, 0.0, "cubic", 30),   # clamp + 3 restarts
    "linear_q4_8": (64, 48, 8, 3, 0.6, 0.4, 0.1, 0.5, "linear", 40),    # clamp + 5 restarts
    "stiff_springs": (64, 48, 16, 1, 0.3, -0.2, 1e4, 0.3, "cubic", 12),  # every trial fails, restarts
    "large_shift": (64, 48, 16, 2, 1.3, -0.8, 0.1, 0.0, "cubic", 40),   # 22 restarts, one fall-through
}


@functools.lru_cache(maxsize=None)
def _problem(name):
    H, side, e, seed, ux, uy, reg, u0s, interp, mi = CASES[name]
    m, mesh, ref, de = structured_case(H=H, W=H, side=side, e=e, seed=seed, ux=ux, uy=uy)
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, ref.shape, roi_mask=m)
    nn = mesh.nodes_xy.shape[0]
    u0 = np.random.default_rng(0).normal(scale=u0s, size=(nn, 2)) if u0s else np.zeros((nn, 2))

    def vg(u):
        return O.objective(u, ref, de, oa, alpha_reg=reg, interp=interp, dtype=np.float64)

    def J(u):
        return O.objective(u, ref, de, oa, alpha_reg=reg, interp=interp, dtype=np.float64, with_grad=False)

    uc, hc, kc = O.cg_solve(u0, ref, de, oa, alpha_reg=reg, max_iter=mi, tol=1e-9, interp=interp, dtype=np.float64)
    return dict(vg=vg, J=J, u0=u0, mi=mi, oracle=(uc, hc, kc), args=(ref, de, oa, reg, interp))


def _agree(a, b):
    (ua, ha, ka), (ub, hb, kb) = a, b
    return (ka == kb and np.array_equal(ha[:, 2], hb[:, 2], equal_nan=True)
            and np.allclose(ha, hb, rtol=1e-9, atol=0, equal_nan=True) and np.allclose(ua, ub, rtol=0, atol=1e-10))


@functools.lru_cache(maxsize=None)
def _restated(name):
    p = _problem(name)
    u, h, k, log = nlcg(p["vg"], p["J"], p["u0"], max_iter=p["mi"], tol=1e-9)
    return u, h, k, log


@pytest.mark.parametrize("name", sorted(CASES))
def test_c_loop_matches_numpy_restatement(name):
    p = _problem(name)
    u, h, k, log = _restated(name)
    uc, hc, kc = p["oracle"]
    assert k == kc == p["mi"]
    np.testing.assert_array_equal(h[:, 2], hc[:, 2])  # identical Armijo decisions
    np.testing.assert_allclose(h, hc, rtol=1e-9, atol=0)
    np.testing.assert_allclose(u, uc, rtol=0, atol=1e-10)
    assert k >= 5


def test_cases_cover_every_branch():
    logs = {n: _restated(n)[3] for n in CASES}
    neg_beta = sum(b is not None and b < 0 for lg in logs.values() for b in lg["beta_raw"])
    restarts = {n: sum(lg["restart"]) for n, lg in logs.items()}
    all_fail = {n: sum(not a for a in lg["accepted"]) for n, lg in logs.items()}
    depths = {t for lg in logs.values() for t in lg["trials"]}
    assert neg_beta >= 2, "no negative Polak-Ribiere ratio: the clamp at 0 is untested"
    assert all(restarts[n] > 0 for n in ("cubic_rigid", "linear_q4_8", "stiff_springs", "large_shift"))
    assert all_fail["stiff_springs"] == CASES["stiff_springs"][-1] and all_fail["large_shift"] >= 1
    assert depths >= {2, 3, 4, 5, 6, 7, 8, 10}, depths
    # the fall-through step is 2^-10 and the history keeps the last trial's J
    h = _restated("stiff_springs")[1]
    assert np.all(h[:, 2] == 2.0 ** -10)


def _beta_fr(g, g_prev, d_prev):
    r = float(np.sum(g * g)) / (float(np.sum(g_prev * g_prev)) + 1e-16)
    return r, r


def _beta_unclamped(g, g_prev, d_prev):
    b, raw = beta_prplus(g, g_prev, d_prev)
    return raw, raw


@pytest.mark.parametrize("mistake,case", [
    (dict(beta_fn=_beta_fr), "cubic_rigid"),
    (dict(beta_fn=_beta_unclamped), "cubic_rigid"),
    (dict(beta_fn=_beta_unclamped), "linear_q4_8"),
    (dict(restart=False), "cubic_rigid"),
    (dict(restart=False), "stiff_springs"),
    (dict(fail_alpha=2.0 ** -9), "stiff_springs"),
    (dict(fail_alpha=0.0), "large_shift"),
])
def test_planted_mistakes_are_caught(mistake, case):
    p = _problem(case)
    u, h, k, _ = nlcg(p["vg"], p["J"], p["u0"], max_iter=p["mi"], tol=1e-9, **mistake)
    assert not _agree((u, h, k), p["oracle"]), f"{mistake} on {case} went unnoticed"


def test_stop_before_update_and_padding():
    """With a tolerance met mid-run the loop records [J, |grad|, 0] at the
    stopping k, leaves u as it was and reports n_iters = k (``:376-389``)."""
    p = _problem("cubic_rigid")
    _, hfull, _, _ = _restated("cubic_rigid")
    tol = float(hfull[7, 1]) * 1.0000001  # the gradient norm seen at k = 7 (or earlier) is below it
    ref, de, oa, reg, interp = p["args"]
    uc, hc, kc = O.cg_solve(p["u0"], ref, de, oa, alpha_reg=reg, max_iter=p["mi"], tol=tol, interp=interp,
                            dtype=np.float64)
    u, h, k, _ = nlcg(p["vg"], p["J"], p["u0"], max_iter=p["mi"], tol=tol)
    assert k == kc <= 7
    assert hc[kc, 2] == 0.0 and hc[kc, 1] < tol and np.isnan(hc[kc + 1:]).all()
    np.testing.assert_allclose(h, hc, rtol=1e-9, atol=0)
    np.testing.assert_allclose(uc, u, rtol=0, atol=1e-10)
    # the stopping evaluation did not move u: it equals the iterate after k steps of the full run
    u_k, _, _, _ = nlcg(p["vg"], p["J"], p["u0"], max_iter=kc, tol=1e-9)
    np.testing.assert_allclose(uc, u_k, rtol=0, atol=1e-10)
