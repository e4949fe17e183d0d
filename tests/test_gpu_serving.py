"""Served sessions: a request solved inside the continuously batched kernel gives
bitwise the same result as the same solve alone (d2ic_cg_solve)."""
from __future__ import annotations

import ctypes
import time

import numpy as np
import pytest

from tests._cases import structured_case

pytestmark = pytest.mark.gpu


def _plan(H, side, e, seed, ux, uy, hole=0.0):
    from d2ic_amd import _lib
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.solver_global_cg import CGPlan

    m, mesh, ref, de = structured_case(H=H, W=H, side=side, e=e, seed=seed, ux=ux, uy=uy, hole=hole)
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
    plan = CGPlan(pa)
    plan.set_reference(ref, "cubic")
    t = _lib.torch()
    return plan, _lib.to_dev(de, t.float32)


def _alone(plan, de, reg, max_iter, tol):
    from d2ic_amd import _lib

    t = _lib.torch()
    u = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
    hist = t.full((max_iter, 3), float("nan"), dtype=t.float32, device=_lib.device())
    cnt = t.zeros(2, dtype=t.int32, device=_lib.device())
    _lib.call("d2ic_cg_solve", plan.handle, _lib.ptr(de), int(de.shape[0]), int(de.shape[1]), _lib.ptr(u),
              float(reg), int(max_iter), float(tol), _lib.ptr(hist), _lib.ptr(cnt),
              ctypes.c_void_p(cnt.data_ptr() + 4), _lib.stream())
    n, p = (int(v) for v in cnt.cpu().tolist())
    return u.cpu().numpy(), hist.cpu().numpy(), n, p


# (H, ROI side, element px, seed, ux, uy, hole radius); the 32 px case has
# elements of 1024 pixels, tiled as 16-column strips
CASES = [(96, 64, 16, 1, 0.3, -0.2, 0.0), (160, 128, 8, 2, -0.4, 0.6, 20.0), (128, 96, 16, 3, 0.8, 0.1, 0.0),
         (160, 128, 32, 4, 0.2, 0.5, 0.0)]


def test_served_equals_alone(gpu):
    from d2ic_amd import _lib
    from d2ic_amd.serving import ServeSession

    t = _lib.torch()
    print("[case] serve: 4 problems, 2 slots", flush=True)
    probs = [_plan(*c) for c in CASES]
    ref = [_alone(plan, de, 0.1, 60, 1e-4) for plan, de in probs]
    sess = ServeSession(slots=2, idle_ms=2.0)
    try:
        for rnd in range(2):
            tickets = [None] * len(probs)
            pending = list(range(len(probs)))
            t0 = time.perf_counter()
            while pending or any(tk is not None and tk.n_iters is None for tk in tickets):
                for i in list(pending):
                    plan, de = probs[i]
                    u = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
                    tk = sess.submit(plan, de, u, 0.1, 60, 1e-4, save_history=True)
                    if tk is not None:
                        tickets[i] = tk
                        pending.remove(i)
                for tk in tickets:
                    if tk is not None and tk.n_iters is None:
                        sess.poll(tk)
                assert time.perf_counter() - t0 < 60
                time.sleep(50e-6)
            for i, tk in enumerate(tickets):
                u_ref, h_ref, n_ref, p_ref = ref[i]
                assert tk.n_iters == n_ref and tk.passes == p_ref, (i, tk.n_iters, n_ref, tk.passes, p_ref)
                assert np.array_equal(tk.u.cpu().numpy(), u_ref), i
                assert np.array_equal(tk.hist.cpu().numpy(), h_ref, equal_nan=True), i
                assert int(tk.counters[0].item()) == n_ref
            print(f"[case] serve round {rnd} ok", flush=True)
            time.sleep(0.05)  # let the kernel leave on idle; the next round relaunches it
    finally:
        sess.close()


def test_stream_server_matches_sequential(gpu):
    """Three streams, three frames each (warm-started), through StreamServer vs the
    plain one-solve-per-launch path."""
    from d2ic_amd import _lib, speckle
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.serving import StreamServer

    t = _lib.torch()
    print("[case] stream server", flush=True)
    streams = []
    for k, (H, e) in enumerate([(96, 16), (128, 8), (112, 16)]):
        m = speckle.rect_mask(H, H, H - 32, H - 32)
        mesh = mask_to_mesh(m, e)
        fns = [speckle.rigid(0.2 * (f + 1), -0.15 * (f + 1)) for f in range(3)]
        ref, frames = speckle.speckle_sequence_device(H, H, 10 + k, fns, noise_sigma=1.0)
        pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
        streams.append((pa, ref, frames))
    srv = StreamServer(slots=4, idle_ms=2.0)
    try:
        for k, (pa, ref, frames) in enumerate(streams):
            assert srv.add_stream(k, pa, ref, max_iter=40, tol=1e-4)
        for f in range(3):
            for k, (_, _, frames) in enumerate(streams):
                srv.submit(k, frames[f])
        srv.drain()
        got = {}
        for r in srv.done:
            got.setdefault(r.key, []).append((r.u.cpu().numpy(), r.n_iters))
        from d2ic_amd.solver_global_cg import CGPlan

        for k, (pa, ref, frames) in enumerate(streams):
            plan = CGPlan(pa)
            plan.set_reference(ref, "cubic")
            u = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
            for f in range(3):
                cnt = t.zeros(2, dtype=t.int32, device=_lib.device())
                _lib.call("d2ic_cg_solve", plan.handle, _lib.ptr(frames[f]), int(frames[f].shape[0]),
                          int(frames[f].shape[1]), _lib.ptr(u), 0.1, 40, 1e-4, None, _lib.ptr(cnt), None,
                          _lib.stream())
                u_s, n_s = got[k][f]
                assert int(cnt[0].item()) == n_s
                assert np.array_equal(u.cpu().numpy(), u_s), (k, f)
            plan.close()
    finally:
        srv.close()


def test_stream_server_device_frame_failures(gpu):
    """A device frame with NaN inside the ROI is found by the served kernel's
    first pass: the request ends REJECTED_INVALID after one pass, its result
    buffer is not used, and the stream's next frame starts from the previous
    warm start and equals the plain solve bitwise.  The previous results stay
    intact (the kernel only reads the warm start)."""
    from d2ic_amd import _lib, speckle
    from d2ic_amd.mask2mesh import mask_to_mesh
    from d2ic_amd.pixel_assets import build_pixel_assets
    from d2ic_amd.serving import DONE, REJECTED_INVALID, StreamServer
    from d2ic_amd.solver_global_cg import CGPlan

    t = _lib.torch()
    print("[case] stream server: NaN device frame", flush=True)
    H, e = 112, 16
    m = speckle.rect_mask(H, H, H - 32, H - 32)
    mesh = mask_to_mesh(m, e)
    fns = [speckle.rigid(0.2 * (f + 1), -0.15 * (f + 1)) for f in range(3)]
    ref, frames = speckle.speckle_sequence_device(H, H, 21, fns, noise_sigma=1.0)
    pa = build_pixel_assets(mesh, ref, 1.0, roi_mask=m)
    bad = frames[1].clone()
    bad[H // 2, H // 2] = float("nan")
    srv = StreamServer(slots=2, idle_ms=2.0)
    try:
        assert srv.add_stream("s", pa, ref, max_iter=40, tol=1e-4)
        r0 = srv.submit("s", frames[0])
        r1 = srv.submit("s", bad)
        r2 = srv.submit("s", frames[2])
        srv.drain()
        assert [r.status for r in (r0, r1, r2)] == [DONE, REJECTED_INVALID, DONE], [r.reason for r in (r0, r1, r2)]
        assert r1.passes == 1 and r1.u is None and "NaN" in r1.reason
        u0_copy = r0.u.clone()
        plan = CGPlan(pa)
        plan.set_reference(ref, "cubic")
        u = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device())
        for f, r in ((0, r0), (2, r2)):
            cnt = t.zeros(2, dtype=t.int32, device=_lib.device())
            _lib.call("d2ic_cg_solve", plan.handle, _lib.ptr(frames[f]), H, H, _lib.ptr(u), 0.1, 40, 1e-4, None,
                      _lib.ptr(cnt), None, _lib.stream())
            assert int(cnt[0].item()) == r.n_iters
            assert t.equal(u, r.u), f
        assert t.equal(r0.u, u0_copy)  # frame 0's result was only read by the later solves
        plan.close()
    finally:
        srv.close()
