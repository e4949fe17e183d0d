"""Adjoint (implicit-function) derivatives of the global solve vs central finite
differences of FP64 oracle solves.

Bars: Hessian-vector products vs central differences of the FP64 oracle
gradient, rtol 1e-3 of max |Hv|; reverse-mode image and reg_strength
gradients at the FP64 solution vs central differences (h = 1e-2 grey,
1e-4 for reg_strength), within 2e-3 of max |FD|; forward tangents likewise.
The reference's own JAX gradients cannot run here (no JAX): parity against
them is unpinned; these checks pin the mathematics."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

H = W = 96
ALPHA = 0.1


def _setup(e=16):
    from d2ic_amd import speckle
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.mesh_assets import build_node_neighbor_tables
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan

    m = speckle.rect_mask(H, W, 64, 64)
    mesh = mask_to_mesh(m, e)
    fn = speckle.affine_field([[2e-3, 1e-3], [-1e-3, 3e-3]], [0.4, -0.3], 0.2, 80.0, W / 2, H / 2)
    ref, de = speckle.speckle_pair(H, W, 21, fn, noise_sigma=1.0)
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
    plan = CGPlan(pa)
    plan.set_reference(ref, "cubic")
    oa = O.pixel_assets(mesh.nodes_xy, mesh.elements, (H, W), roi_mask=m)
    nb_idx, nb_deg = build_node_neighbor_tables(mesh)
    return m, mesh, ref, de, plan, oa, np.asarray(nb_idx), np.asarray(nb_deg)


def _u64(ref, de, oa, u0, alpha=ALPHA, newton=4):
    """Tight FP64 solution: oracle NLCG, then Newton steps whose linear systems
    are solved by CG with Hessian-vector products from central differences of
    the oracle's FP64 gradient (NLCG alone stalls near |grad J| ~ 1e-8)."""
    u, _, _ = O.cg_solve(u0, ref, de, oa, alpha_reg=alpha, max_iter=300, tol=1e-10, interp="cubic",
                         dtype=np.float64)
    act = (oa["node_pixel_degree"] > 0)[:, None]

    def grad(x):
        return O.objective(x, ref, de, oa, alpha_reg=alpha, interp="cubic", dtype=np.float64)[1] * act

    for _ in range(newton):
        g = grad(u)
        if np.abs(g).max() < 1e-15:
            break

        def hv(v):
            e = 1e-4 / max(np.abs(v).max(), 1e-30)
            return (grad(u + e * v) - grad(u - e * v)) / (2 * e)

        x = np.zeros_like(u)
        r = -g
        p = r.copy()
        rr = (r * r).sum()
        for _ in range(200):
            hp = hv(p)
            a = rr / (p * hp).sum()
            x += a * p
            r -= a * hp
            rn = (r * r).sum()
            if rn < 1e-30 * (g * g).sum() + 1e-60:
                break
            p = r + (rn / rr) * p
            rr = rn
        u = u + x
    return u


def _exx_np(u, X, idx, deg):
    out = np.zeros(len(X))
    for i in range(len(X)):
        d = int(deg[i])
        if d == 0:
            continue
        dX = X[idx[i, :d]] - X[i]
        dU = u[idx[i, :d]] - u[i]
        A = dX.T @ dX + 1e-8 * np.eye(2)
        G = np.linalg.solve(A, dX.T @ dU).T  # G[c, j] = d u_c / d x_j
        F = np.eye(2) + G
        out[i] = 0.5 * (F[:, 0] @ F[:, 0] - 1.0)
    return out


@pytest.mark.parametrize("e", [16, 32])
def test_hvp_matches_gradient_differences(gpu, e):
    """e = 32: elements of 1024 pixels, whose tiles are 16-column strips."""
    from d2ic_amd.adjoint import Linearization

    print(f"[case] adjoint hvp e={e}", flush=True)
    m, mesh, ref, de, plan, oa, _, _ = _setup(e)
    t = gpu
    u0 = np.zeros((plan.n_nodes, 2))
    u = _u64(ref, de, oa, u0)
    lin = Linearization(plan, de, u.astype(np.float32))
    rng = np.random.default_rng(0)
    act = plan.node_active.cpu().numpy()
    v = rng.normal(size=u.shape) * act[:, None]
    hv = lin.hvp(t.tensor(v[None], dtype=t.float32, device="cuda"), ALPHA)[0].cpu().numpy()
    eps = 1e-4
    _, gp = O.objective(u.astype(np.float32) + eps * v, ref, de, oa, alpha_reg=ALPHA, interp="cubic",
                        dtype=np.float64)
    _, gm = O.objective(u.astype(np.float32) - eps * v, ref, de, oa, alpha_reg=ALPHA, interp="cubic",
                        dtype=np.float64)
    fd = (gp - gm) / (2 * eps)
    err = np.abs(hv - fd)[act].max() / np.abs(fd[act]).max()
    print(f"hvp rel err {err:.2e}")
    assert err < 1e-3
    w = rng.normal(size=u.shape) * act[:, None]
    hw = lin.hvp(t.tensor(w[None], dtype=t.float32, device="cuda"), ALPHA)[0].cpu().numpy()
    assert abs((v * hw).sum() - (w * hv).sum()) < 1e-4 * abs((v * hw).sum())


def test_reverse_mode_matches_finite_differences(gpu):
    from d2ic_amd.adjoint import Linearization, green_lagrange_torch

    print("[case] adjoint reverse mode", flush=True)
    m, mesh, ref, de, plan, oa, nb_idx, nb_deg = _setup()
    t = gpu
    X = np.asarray(mesh.nodes_xy, np.float64)
    act = plan.node_active.cpu().numpy()
    win = act & (np.abs(X[:, 0] - W / 2) <= 16) & (np.abs(X[:, 1] - H / 2) <= 16)
    rng = np.random.default_rng(1)
    u_t = rng.normal(scale=0.2, size=(plan.n_nodes, 2)) * act[:, None]

    def loss_np(u):
        return float(((u - u_t)[act] ** 2).sum() + _exx_np(u, X, nb_idx, nb_deg)[win].mean())

    u64 = _u64(ref, de, oa, np.zeros((plan.n_nodes, 2)))
    # dL/du with torch autograd on the differentiable strain
    ut = t.tensor(u64, dtype=t.float64, device="cuda", requires_grad=True)
    act_t = t.tensor(act, device="cuda")
    win_t = t.tensor(win, device="cuda")
    E = green_lagrange_torch(ut, X, nb_idx, nb_deg)
    L = ((ut - t.tensor(u_t, device="cuda"))[act_t] ** 2).sum() + E[win_t, 0, 0].mean()
    L.backward()
    dLdu = ut.grad.float()
    lin = Linearization(plan, de, u64.astype(np.float32))
    lam = lin.solve(dLdu[None], ALPHA, rtol=1e-9)[0]
    gf, gg = lin.image_grads(lam)
    gf, gg = gf.cpu().numpy(), gg.cpu().numpy()
    galpha = float(-(lam.double() * lin.spring(t.tensor(u64[None], dtype=t.float32, device="cuda"))[0].double()).sum())
    rows, cols = np.nonzero(m[20:76, 20:76])
    pick = rng.choice(len(rows), size=6, replace=False)
    h = 1e-2
    for img_name in ("g", "f"):
        fd, ad = [], []
        for k in pick:
            r, c = rows[k] + 20, cols[k] + 20
            imgs = []
            for sgn in (1, -1):
                rr, dd = ref.copy(), de.copy()
                if img_name == "g":
                    dd[r, c] += sgn * h
                else:
                    rr[r, c] += sgn * h
                imgs.append(loss_np(_u64(rr, dd, oa, u64)))
            fd.append((imgs[0] - imgs[1]) / (2 * h))
            ad.append((gg if img_name == "g" else gf)[r, c])
        fd, ad = np.array(fd), np.array(ad)
        err = np.abs(fd - ad).max() / np.abs(fd).max()
        print(f"d/d{img_name}: fd {fd} adj {ad} rel err {err:.2e}")
        assert err < 2e-3
    ha = 1e-4
    fa = (loss_np(_u64(ref, de, oa, u64, ALPHA + ha)) - loss_np(_u64(ref, de, oa, u64, ALPHA - ha))) / (2 * ha)
    print(f"d/dalpha: fd {fa:.6e} adj {galpha:.6e}")
    assert abs(fa - galpha) < 2e-3 * abs(fa)


def test_forward_tangent_matches_finite_differences(gpu):
    from d2ic_amd.adjoint import Linearization

    print("[case] adjoint forward tangent", flush=True)
    m, mesh, ref, de, plan, oa, _, _ = _setup()
    t = gpu
    act = plan.node_active.cpu().numpy()
    u64 = _u64(ref, de, oa, np.zeros((plan.n_nodes, 2)))
    lin = Linearization(plan, de, u64.astype(np.float32))
    rng = np.random.default_rng(3)
    dg = rng.normal(size=(H, W)).astype(np.float32)
    du = lin.tangent(t.tensor(dg[None], device="cuda"), ALPHA, "g", rtol=1e-9)[0].cpu().numpy()
    eps = 1e-2
    fd = (_u64(ref, de + eps * dg, oa, u64) - _u64(ref, de - eps * dg, oa, u64)) / (2 * eps)
    err = np.abs(du - fd)[act].max() / np.abs(fd[act]).max()
    print(f"tangent rel err {err:.2e}")
    assert err < 2e-3
    assert np.all(du[~act] == 0.0)


def test_block_jacobi_preconditioner(gpu):
    """d2ic_adj_diag_inv against the 2x2 diagonal blocks of H read off
    Hessian-vector products with unit vectors (inverted in FP64); the
    preconditioned and plain CG solves agree."""
    from d2ic_amd.adjoint import Linearization

    print("[case] adjoint block-Jacobi preconditioner", flush=True)
    m, mesh, ref, de, plan, oa, _, _ = _setup()
    t = gpu
    u = _u64(ref, de, oa, np.zeros((plan.n_nodes, 2)), newton=1)
    lin = Linearization(plan, de, u.astype(np.float32))
    nn = plan.n_nodes
    act = plan.node_active.cpu().numpy()
    E = np.zeros((2 * nn, nn, 2), np.float32)
    for n in range(nn):
        E[2 * n, n, 0] = 1.0
        E[2 * n + 1, n, 1] = 1.0
    HE = lin.hvp(t.tensor(E, device="cuda"), ALPHA).cpu().numpy().astype(np.float64)
    A, C = (x.cpu().numpy() for x in lin.precond(ALPHA))
    checked = 0
    for n in np.nonzero(act)[0]:
        B = np.array([[HE[2 * n, n, 0], HE[2 * n + 1, n, 0]], [HE[2 * n, n, 1], HE[2 * n + 1, n, 1]]])
        B = 0.5 * (B + B.T)
        if B[0, 0] <= 0 or np.linalg.det(B) <= 1e-12 * (abs(B[0, 0]) + abs(B[1, 1])) ** 2:
            continue
        inv = np.linalg.inv(B)
        got = np.array([[A[n, 0], C[n, 0]], [C[n, 1], A[n, 1]]])
        assert np.abs(got - inv).max() <= 1e-4 * np.abs(inv).max(), (n, got, inv)
        checked += 1
    assert checked >= act.sum() // 2
    rng = np.random.default_rng(1)
    b = t.tensor((rng.normal(size=(1, nn, 2)) * act[None, :, None]), dtype=t.float32, device="cuda")
    x0 = lin.solve(b, ALPHA, rtol=1e-7, precondition=False)
    it0 = lin.last_iters
    x1 = lin.solve(b, ALPHA, rtol=1e-7, precondition=True)
    it1 = lin.last_iters
    rel = float((x1 - x0).norm() / x0.norm())
    print(f"checked {checked} blocks; iterations plain {it0}, preconditioned {it1}; solution rel diff {rel:.2e}")
    assert rel < 1e-4


def test_graph_captured_cg_equals_eager(gpu):
    """The linear CG's HIP-graph replays (check_every iterations per replay)
    give bitwise the eager iterates, over several replays per solve and
    several solves with eager work and synchronisations in between."""
    from d2ic_amd.adjoint import Linearization

    print("[case] adjoint CG: graph replays vs eager", flush=True)
    m, mesh, ref, de, plan, oa, _, _ = _setup()
    t = gpu
    u = _u64(ref, de, oa, np.zeros((plan.n_nodes, 2)), newton=1)
    lin = Linearization(plan, de, u.astype(np.float32))
    act = plan.node_active.cpu().numpy()
    rng = np.random.default_rng(5)
    b = t.tensor(rng.normal(size=(3, plan.n_nodes, 2)) * act[None, :, None], dtype=t.float32, device="cuda")
    for pre, every in ((False, 10), (True, 5)):
        xe = lin.solve(b, ALPHA, rtol=1e-8, precondition=pre, check_every=every, graph=False)
        ite = lin.last_iters
        assert ite >= 4 * every  # at least three replays after the eager block
        for rep in range(3):
            xg = lin.solve(b, ALPHA, rtol=1e-8, precondition=pre, check_every=every, graph=True, graph_after=1)
            print(f"precondition={pre} check_every={every} rep {rep}: iterations {lin.last_iters} (eager {ite})")
            assert lin.last_iters == ite
            assert t.equal(xg, xe)
            _ = lin.hvp(b, ALPHA)  # eager work between the graph solves
            t.cuda.synchronize()
