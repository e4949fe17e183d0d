"""Differentiation of global solves through hand-written adjoint kernels.

The reference differentiates through ``_cg_solve`` with ``jax.grad`` /
``jax.jvp`` (uncertainty propagation; reference notebook 11).  Here the
solution u* of G(u) = grad_u J(u; f, g, alpha) = 0 (``solver_global_cg.py:
289-323``) is differentiated implicitly:

* reverse mode: dL/dtheta = -(dG/dtheta)^T lambda with H lambda = dL/du
  (one linear CG solve with Hessian-vector products, ``d2ic_adj_hvp``);
* forward mode: du = H^-1 (-(dG/dtheta) dtheta) for perturbation images
  (``d2ic_adj_tangent``), batched over many probes.

H is the exact Hessian of J at u* (including the r * hess(g) term) unless
``gauss_newton=True``.  Nodes that no pixel and no spring touches keep u0 in
the solver; their derivatives are zero and they are masked out of the linear
solves.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .solver_global_cg import CGPlan


class Linearization:
    """The image term linearised at u for one deformed image."""

    def __init__(self, plan: CGPlan, def_image, u, gauss_newton: bool = False):
        t = _lib.require_gpu()
        self.plan = plan
        self.g = _lib.to_dev(def_image, t.float32)
        self.H, self.W = int(self.g.shape[0]), int(self.g.shape[1])
        self.u = _lib.to_dev(u, t.float32, (plan.n_nodes, 2)).contiguous()
        a, b, c = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.call("d2ic_adj_sizes", plan.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.tile_floats = int(c.value)
        dev = _lib.device()
        self.lin = t.empty(max(int(a.value), 1), dtype=t.float32, device=dev)
        self.K = t.empty(max(int(b.value), 1), dtype=t.float32, device=dev)
        self.mask = plan.node_active.to(t.float32)[:, None]
        _lib.call("d2ic_adj_linearize", plan.handle, _lib.ptr(self.g), self.H, self.W, _lib.ptr(self.u),
                  int(bool(gauss_newton)), _lib.ptr(self.lin), _lib.ptr(self.K), _lib.stream())

    def _work(self, nv):
        t = _lib.torch()
        return t.empty(max(nv * self.tile_floats, 1), dtype=t.float32, device=_lib.device())

    def hvp(self, V, alpha):
        """H v for V of shape (nv, n_nodes, 2)."""
        t = _lib.torch()
        V = V.contiguous()
        out = t.empty_like(V)
        nv = int(V.shape[0])
        _lib.call("d2ic_adj_hvp", self.plan.handle, _lib.ptr(self.K), float(alpha), _lib.ptr(V), nv,
                  _lib.ptr(out), _lib.ptr(self._work(nv)), _lib.stream())
        return out

    def spring(self, V):
        """L_W v (spring-energy gradient operator, without alpha)."""
        t = _lib.torch()
        V = V.contiguous()
        out = t.empty_like(V)
        nv = int(V.shape[0])
        _lib.call("d2ic_adj_spring", self.plan.handle, _lib.ptr(V), nv, _lib.ptr(out), _lib.ptr(self._work(nv)),
                  _lib.stream())
        return out

    def precond(self, alpha):
        """Block-Jacobi preconditioner: the inverse 2x2 diagonal block of H per
        node as two (n_nodes, 2) factors (A, C) with M^-1 r = A * r + C * flip(r)."""
        t = _lib.torch()
        if getattr(self, "_pc_alpha", None) != float(alpha):
            inv = t.empty((self.plan.n_nodes, 3), dtype=t.float32, device=_lib.device())
            _lib.call("d2ic_adj_diag_inv", self.plan.handle, _lib.ptr(self.K), float(alpha), _lib.ptr(inv),
                      _lib.stream())
            inv = inv * self.mask
            self._pc = (t.stack([inv[:, 0], inv[:, 2]], 1).contiguous(), t.stack([inv[:, 1], inv[:, 1]], 1).contiguous())
            self._pc_alpha = float(alpha)
        return self._pc

    def solve(self, B, alpha, rtol: float = 1e-6, max_iter: int = 2000, check_every: int = 10,
              precondition: bool = True, graph: bool = True, graph_after: int = 3):
        """Conjugate gradients on H X = B for every right-hand side of B (nv, n_nodes, 2),
        block-Jacobi preconditioned unless ``precondition=False``; FP64 dot
        products, per-RHS step sizes; stops when every |R| <= rtol |B|
        (tested every ``check_every`` iterations).

        With ``graph=True`` a solve still running after ``graph_after``
        blocks of ``check_every`` eager iterations continues as replays of
        one HIP graph holding ``check_every`` iterations (the ~15 launches of
        an iteration cost more than their work on meshes of a few thousand
        nodes; capturing costs about one block, so short solves stay eager).
        The graph runs the same kernels in the same order on the same
        buffers, so the iterates are bitwise those of ``graph=False``."""
        t = _lib.torch()
        m = self.mask
        B = B * m
        X = t.zeros_like(B)
        R = B.clone()
        if precondition:
            A, C = self.precond(alpha)

            def apply(V):
                return t.addcmul(A * V, C, V.flip(-1))
        else:
            def apply(V):
                return V.clone()
        P = apply(R)
        rz = (R.double() * P.double()).sum(dim=(1, 2))
        stop = (rtol ** 2) * (R.double() ** 2).sum(dim=(1, 2)).clamp_min(1e-300)

        def step(update_p: bool):
            HP = self.hvp(P, alpha) * m
            php = (P.double() * HP.double()).sum(dim=(1, 2))
            a = t.where(php > 0, rz / php.clamp_min(1e-300), t.zeros_like(rz)).float()[:, None, None]
            X.addcmul_(a, P)
            R.addcmul_(a, HP, value=-1.0)
            if update_p:
                Z = apply(R)
                rz_new = (R.double() * Z.double()).sum(dim=(1, 2))
                beta = t.where(rz > 0, rz_new / rz.clamp_min(1e-300), t.zeros_like(rz)).float()[:, None, None]
                P.copy_(t.addcmul(Z, beta, P))
                rz.copy_(rz_new)

        def converged():
            return bool(((R.double() ** 2).sum(dim=(1, 2)) <= stop).all())

        k = max(int(check_every), 1)
        it = 0
        g = None
        while it < max_iter:
            n = min(k, max_iter - it)
            if graph and g is None and it >= max(int(graph_after), 1) * k and n == k:
                g = t.cuda.CUDAGraph()
                with t.cuda.graph(g):
                    for _ in range(k):
                        step(True)
            if g is not None and n == k:
                g.replay()
            else:
                for _ in range(n):
                    step(True)
            it += n
            if n == k and converged():
                break
        self.last_iters = it
        return X

    def tangent_rhs(self, dimg, which: str = "g"):
        """-(dG/dg) dg (which='g') or -(dG/df) df (which='f') for images (nv, H, W)."""
        t = _lib.torch()
        D = _lib.to_dev(dimg, t.float32)
        if D.dim() == 2:
            D = D[None]
        nv = int(D.shape[0])
        B = t.empty((nv, self.plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
        _lib.call("d2ic_adj_tangent", self.plan.handle, self.H, self.W, _lib.ptr(self.u), _lib.ptr(self.lin),
                  0 if which == "g" else 1, _lib.ptr(D.contiguous()), nv, _lib.ptr(B), _lib.ptr(self._work(nv)),
                  _lib.stream())
        return B

    def tangent(self, dimg, alpha, which: str = "g", **cg):
        """Forward-mode derivative du of the solution for perturbation images (nv, H, W)."""
        return self.solve(self.tangent_rhs(dimg, which), alpha, **cg)

    def image_grads(self, lam, want_f: bool = True, want_g: bool = True):
        """(dL/df, dL/dg) images for the adjoint state lambda (n_nodes, 2)."""
        t = _lib.torch()
        lam = _lib.to_dev(lam, t.float32, (self.plan.n_nodes, 2)).contiguous()
        dev = _lib.device()
        gf = t.empty((self.H, self.W), dtype=t.float32, device=dev) if want_f else None
        gg = t.empty((self.H, self.W), dtype=t.float32, device=dev) if want_g else None
        work = t.empty((self.H * self.W,), dtype=t.int64, device=dev) if want_g else None
        _lib.call("d2ic_adj_image_grads", self.plan.handle, self.H, self.W, _lib.ptr(self.u), _lib.ptr(self.lin),
                  _lib.ptr(lam), _lib.ptr(gf), _lib.ptr(gg), _lib.ptr(work), _lib.stream())
        return gf, gg


def _solve(plan: CGPlan, def_image, u0, alpha, max_iter, tol):
    t = _lib.torch()
    g = _lib.to_dev(def_image, t.float32)
    u = _lib.to_dev(u0, t.float32, (plan.n_nodes, 2)).clone()
    cnt = t.zeros(2, dtype=t.int32, device=_lib.device())
    _lib.call("d2ic_cg_solve", plan.handle, _lib.ptr(g), int(g.shape[0]), int(g.shape[1]), _lib.ptr(u), float(alpha),
              int(max_iter), float(tol), None, _lib.ptr(cnt), ctypes.c_void_p(cnt.data_ptr() + 4), _lib.stream())
    return u, int(cnt[0].item())


def _make_fn():
    torch = _lib.torch()

    class ImplicitGlobalSolve(torch.autograd.Function):
        """u* = argmin_u J(u; f, g, alpha) with implicit-function gradients."""

        @staticmethod
        def forward(ctx, ref_image, def_image, reg_strength, plan, u0, max_iter, tol, interpolation, cg_opts):
            plan.set_reference(ref_image.detach(), interpolation)
            alpha = float(reg_strength.detach().item())
            u, n = _solve(plan, def_image.detach(), u0, alpha, max_iter, tol)
            ctx.plan, ctx.alpha, ctx.cg_opts = plan, alpha, cg_opts
            ctx.save_for_backward(def_image.detach(), u)
            ctx.n_iters = n
            return u

        @staticmethod
        def backward(ctx, grad_u):
            g, u = ctx.saved_tensors
            lin = Linearization(ctx.plan, g, u, gauss_newton=ctx.cg_opts.get("gauss_newton", False))
            opts = {k: v for k, v in ctx.cg_opts.items() if k != "gauss_newton"}
            lam = lin.solve(grad_u.float().contiguous()[None], ctx.alpha, **opts)[0]
            gf, gg = lin.image_grads(lam, ctx.needs_input_grad[0], ctx.needs_input_grad[1])
            galpha = None
            if ctx.needs_input_grad[2]:
                galpha = -(lam.double() * lin.spring(u[None])[0].double()).sum().to(torch.float32)
            return gf, gg, galpha, None, None, None, None, None, None

    return ImplicitGlobalSolve


_FN = None


def solve_differentiable(plan: CGPlan, ref_image, def_image, reg_strength, u0=None, *, max_iter: int = 200,
                         tol: float = 1e-3, interpolation: str = "cubic", gauss_newton: bool = False,
                         cg_rtol: float = 1e-6, cg_max_iter: int = 2000):
    """Global solve whose output carries gradients w.r.t. both images and
    reg_strength (torch tensors with requires_grad)."""
    global _FN
    if _FN is None:
        _FN = _make_fn()
    t = _lib.torch()
    if not isinstance(reg_strength, t.Tensor):
        reg_strength = t.tensor(float(reg_strength), device=_lib.device())
    if u0 is None:
        u0 = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
    opts = dict(gauss_newton=gauss_newton, rtol=cg_rtol, max_iter=cg_max_iter)
    return _FN.apply(ref_image, def_image, reg_strength, plan, u0, max_iter, tol, interpolation, opts)


_ROW_GATHER = None


def _row_gather_fn():
    """u[..., idx, :] whose backward is a fixed-order gather over the
    transpose of idx (torch's own backward of that indexing accumulates the
    repeated node indices serially and took ~6 ms for 4,225 nodes)."""
    t = _lib.torch()

    class RowGather(t.autograd.Function):
        @staticmethod
        def forward(ctx, u, flat_idx, tpos):
            # u has one row more than tpos (the zero row that padding slots read)
            ctx.save_for_backward(tpos)
            return u.index_select(-2, flat_idx)

        @staticmethod
        def backward(ctx, g):
            (tpos,) = ctx.saved_tensors
            n, width = int(tpos.shape[0]), int(tpos.shape[1])
            gz = t.cat([g, t.zeros_like(g[..., :1, :])], dim=-2)  # position m reads zero
            out = gz.index_select(-2, tpos.reshape(-1)).reshape(*g.shape[:-2], n, width, g.shape[-1]).sum(-2)
            return t.cat([out, t.zeros_like(out[..., :1, :])], dim=-2), None, None

    return RowGather


def _gather_tables(idx, valid, n_rows):
    """flat row index per (node, slot) (padding -> the appended zero row
    n_rows) and the transpose table: for each row, the flat positions that
    read it in ascending order, padded with the zero-row position."""
    t = _lib.torch()
    flat = t.where(valid, idx, t.full_like(idx, n_rows)).reshape(-1)
    m = flat.numel()
    keys, pos = t.sort(flat, stable=True)
    cnt = t.bincount(flat, minlength=n_rows + 1)[:n_rows]
    start = t.cumsum(cnt, 0) - cnt
    real = keys < n_rows
    keys, pos = keys[real], pos[real]
    rank = t.arange(keys.numel(), device=idx.device) - start[keys]
    width = max(int(cnt.max().item()) if n_rows > 0 else 0, 1)
    tpos = t.full((n_rows, width), m, dtype=t.int64, device=idx.device)
    tpos[keys, rank] = pos
    return flat, tpos


def green_lagrange_torch(u, nodes, nb_index, nb_degree, gauge_length: float = 0.0, eps: float = 1e-8):
    """Nodal LSQ Green-Lagrange tensor (``strain.py:11-114``) as differentiable
    torch code: returns E (..., n_nodes, 2, 2) for u (..., n_nodes, 2)."""
    t = _lib.torch()
    X = _lib.to_dev(nodes, t.float32).reshape(-1, 2).to(u.dtype)
    idx = _lib.to_dev(nb_index, t.int64)
    deg = _lib.to_dev(nb_degree, t.int64).reshape(-1)
    md = idx.shape[1]
    valid = (t.arange(md, device=idx.device)[None, :] < deg[:, None])
    dX = X[idx] - X[:, None, :]
    global _ROW_GATHER
    if _ROW_GATHER is None:
        _ROW_GATHER = _row_gather_fn()
    nn = int(u.shape[-2])
    flat, tpos = _gather_tables(idx, valid, nn)
    u_ext = t.cat([u, t.zeros_like(u[..., :1, :])], dim=-2)
    uj = _ROW_GATHER.apply(u_ext, flat, tpos).reshape(*u.shape[:-2], nn, md, u.shape[-1])
    dU = uj - u[..., :, None, :]
    if gauge_length > 0:
        w = t.exp(-((dX ** 2).sum(-1)) / (gauge_length ** 2))
    else:
        w = t.ones(dX.shape[:2], dtype=u.dtype, device=u.device)
    w = w * valid.to(u.dtype)
    A = t.einsum("nk,nki,nkj->nij", w, dX, dX) + eps * t.eye(2, dtype=u.dtype, device=u.device)
    Bm = t.einsum("nk,...nki,nkj->...nij", w, dU, dX)  # rows: component of u, cols: direction
    # grad_u[r] = A^-1 Bm[r] for each component r (the reference's two 2x2
    # solves), by Cramer's rule: elementwise, so its backward is elementwise too
    a, b, c, d = A[:, 0, 0], A[:, 0, 1], A[:, 1, 0], A[:, 1, 1]
    det = a * d - b * c
    B0, B1 = Bm[..., 0], Bm[..., 1]
    grad_u = t.stack([(d[:, None] * B0 - b[:, None] * B1), (a[:, None] * B1 - c[:, None] * B0)], -1) / det[:, None, None]
    F = t.eye(2, dtype=u.dtype, device=u.device) + grad_u
    return 0.5 * (F.transpose(-1, -2) @ F - t.eye(2, dtype=u.dtype, device=u.device))


def strain_noise_std(lin: Linearization, alpha, n_probes: int = 256, sigma: float = 1.0, seed: int = 0,
                     batch: int = 64, strain_fn=None, **cg):
    """Noise-to-displacement (and optionally strain) standard deviations from
    forward-tangent probes with i.i.d. N(0, sigma^2) noise on the deformed image."""
    t = _lib.torch()
    gen = t.Generator(device=_lib.device())
    gen.manual_seed(seed)
    du_sq = t.zeros((lin.plan.n_nodes, 2), dtype=t.float64, device=_lib.device())
    e_sq = None
    done = 0
    while done < n_probes:
        nv = min(batch, n_probes - done)
        noise = sigma * t.randn((nv, lin.H, lin.W), generator=gen, device=_lib.device(), dtype=t.float32)
        du = lin.tangent(noise, alpha, "g", **cg)
        du_sq += (du.double() ** 2).sum(0)
        if strain_fn is not None:
            e = t.stack([strain_fn(du[k]) for k in range(nv)]).double()
            e_sq = (e ** 2).sum(0) if e_sq is None else e_sq + (e ** 2).sum(0)
        done += nv
    out = dict(u_std=t.sqrt(du_sq / n_probes).float())
    if e_sq is not None:
        out["strain_std"] = t.sqrt(e_sq / n_probes).float()
    return out
