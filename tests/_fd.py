"""Finite-difference machinery on the FP64 oracle, shared by the derivative
tests: Hessian-vector products from central differences of the oracle's
gradient, a linear CG on them, and a Newton polish of an NLCG solution (NLCG
alone stalls near |grad J| ~ 1e-8)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def oracle_grad(ref, de, oa, alpha, act):
    """u -> grad J(u) (FP64 oracle), zero on nodes that nothing moves."""
    def grad(x):
        return O.objective(x, ref, de, oa, alpha_reg=alpha, interp="cubic", dtype=np.float64)[1] * act
    return grad


def fd_hvp(grad, u, v, rel=1e-4):
    e = rel / max(float(np.abs(v).max()), 1e-30)
    return (grad(u + e * v) - grad(u - e * v)) / (2 * e)


def cg_fd(grad, u, b, rtol=1e-12, max_iter=500):
    """Solve H x = b with H the FD Hessian of grad at u (plain CG)."""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = float((r * r).sum())
    b2 = float((b * b).sum())
    for _ in range(max_iter):
        if rr <= rtol * rtol * b2 + 1e-300:
            break
        hp = fd_hvp(grad, u, p)
        a = rr / float((p * hp).sum())
        x += a * p
        r -= a * hp
        rn = float((r * r).sum())
        p = r + (rn / rr) * p
        rr = rn
    return x


def tight_solution(ref, de, oa, u0, alpha, act, nlcg_iters=300, newton=4):
    """FP64 minimiser: oracle NLCG from u0, then Newton steps solved by CG on
    FD Hessian-vector products."""
    u, _, _ = O.cg_solve(u0, ref, de, oa, alpha_reg=alpha, max_iter=nlcg_iters, tol=1e-10, interp="cubic",
                         dtype=np.float64)
    grad = oracle_grad(ref, de, oa, alpha, act)
    for _ in range(newton):
        g = grad(u)
        if np.abs(g).max() < 1e-15:
            break
        u = u + cg_fd(grad, u, -g, rtol=1e-12, max_iter=400)
    return u
