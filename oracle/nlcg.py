"""Second, independent restatement of the reference's NLCG loop -- TEST
INFRASTRUCTURE ONLY (a checker of the C oracle's ``orc_cg_solve``; only
``tests/`` import it).

``nlcg`` follows ``_cg_solve`` (``d2ic/solver_global_cg.py:326-469``) line by
line in numpy, but takes the objective as two callables, so it can be driven
by the C oracle's FD-pinned ``objective`` and compared with the C loop
iteration by iteration:

* ``while k < max_iter and not done`` (``:364-367``); value and gradient at u
  (``:371-375``); ``|grad|_2 < tol`` stops before the update, records
  ``[J, |grad|, 0]`` and leaves k unchanged (``:376-377, 379-389``);
* d = -g at k = 0, else -g + max(0, g.(g - g_prev) / (g_prev.g_prev + 1e-16)) d
  (``:391-404``); restart d = -g with g.d = -|g|^2 when g.d >= 0
  (``:405-413``);
* Armijo: alpha = 1, 10 trials, accept the first with
  J(u + alpha d) <= J + 1e-4 alpha g.d, else halve; the record keeps the last
  evaluated J; after 10 failures alpha = 2^-10, applied without evaluation
  (``:415-448``); u += alpha d, ``hist[k] = [J_trial, |grad|, alpha]``
  (``:457-459``).

It also returns a log of the decisions (beta before the clamp, restarts,
line-search trial counts) so the tests can show which branches a case hit.
``beta_fn``, ``restart`` and ``fail_alpha`` exist so a test can plant a wrong
rule here and check that the comparison with the C oracle catches it.
"""
from __future__ import annotations

import numpy as np


def beta_prplus(g, g_prev, d_prev):
    """Polak-Ribiere-plus (``solver_global_cg.py:394-396``); also returns the raw ratio."""
    num = float(np.sum(g * (g - g_prev)))
    den = float(np.sum(g_prev * g_prev)) + 1e-16
    raw = num / den
    return max(raw, 0.0), raw


def nlcg(vg, J, u0, max_iter=200, tol=1e-3, beta_fn=beta_prplus, restart=True, fail_alpha=None, c_armijo=1e-4,
         trials=10):
    """Returns (u, hist (max_iter, 3) NaN padded, n_iters, log) in float64.

    ``vg(u) -> (J, grad)`` and ``J(u) -> J`` act on (Nn, 2) float64 arrays.
    ``fail_alpha`` (None: the reference's 2^-10 left by the halvings)
    replaces the step applied when every trial fails."""
    u = np.array(u0, dtype=np.float64).copy()
    hist = np.full((int(max_iter), 3), np.nan)
    d = np.zeros_like(u)
    g_prev = np.zeros_like(u)
    log = dict(beta_raw=[], restart=[], trials=[], accepted=[])
    k = 0
    done = False
    while k < max_iter and not done:
        Jv, g = vg(u)
        g = np.asarray(g, np.float64)
        gnorm = float(np.sqrt(np.sum(g * g)))
        if gnorm < tol:
            hist[k] = (Jv, gnorm, 0.0)
            done = True
            break
        if k == 0:
            dn = -g
            log["beta_raw"].append(None)
        else:
            beta, raw = beta_fn(g, g_prev, d)
            log["beta_raw"].append(raw)
            dn = -g + beta * d
        gd = float(np.sum(g * dn))
        restarted = restart and gd >= 0
        if restarted:
            dn = -g
            gd = -float(np.sum(g * g))
        log["restart"].append(bool(restarted))
        alpha, J_trial, ok = 1.0, Jv, False
        n_trials = 0
        for _ in range(trials):
            if ok:
                break
            J_trial = J(u + alpha * dn)
            n_trials += 1
            ok = J_trial <= Jv + c_armijo * alpha * gd
            if not ok:
                alpha *= 0.5
        if not ok and fail_alpha is not None:
            alpha = float(fail_alpha)
        log["trials"].append(n_trials)
        log["accepted"].append(bool(ok))
        u = u + alpha * dn
        hist[k] = (J_trial, gnorm, alpha)
        d, g_prev = dn, g
        k += 1
    return u, hist, k, log
