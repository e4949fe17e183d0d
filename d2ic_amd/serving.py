"""Serving: continuous batching of global CG solves in one persistent kernel.

No counterpart in the reference (which solves one frame per call,
``solver_global_cg.py:155-198``).  Requests (frame pairs of many camera
streams, each with its own mesh/ROI) join the running persistent CG kernel at
its next grid barrier and leave when solved, so solves of different sizes
overlap on the device.  Each request runs exactly the computation of
``GlobalCGSolver.solve`` and gives bitwise the same nodal displacements.

* :class:`ServeSession` -- thin wrapper of ``d2ic_serve_*`` (slots, tickets).
* :class:`StreamServer` -- per-stream resident state (plan with f sampled at the
  pixels, warm start), admission of streams against an HBM budget, bounded
  per-stream queues, per-request status codes, cancellation, per-stream frame
  ordering (frame k+1 starts from frame k's solution), latency records.

Request status codes (``Request.status``):

=====================  =========================================================
``QUEUED``             waiting in its stream's queue
``RUNNING``            submitted to the served kernel
``DONE``               solved; ``u``, ``n_iters``, ``passes`` are set
``REJECTED_INVALID``   frame not 2-D or of another shape than the stream's
                       reference image, or a host frame holding NaN/Inf: never
                       submitted; or a device frame (or warm start) whose
                       NaN/Inf values the solve samples: found by the served
                       kernel in its first pass, which ends the request
``REJECTED_QUEUE``     the stream's queue already holds ``max_queue`` frames
``REJECTED_STREAM``    unknown (or removed) stream key
``CANCELLED``          cancelled while queued, or while running (the result is
                       discarded and does not become the stream's warm start)
``FAILED``             the solve did not finish (a grid barrier timed out) or
                       ended with a non-finite objective or gradient (reported
                       by the kernel); the stream keeps its previous warm start
=====================  =========================================================
"""
from __future__ import annotations

import ctypes
import sys
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from . import _lib
from .solver_global_cg import CGPlan, _normalize_interpolation

QUEUED = "QUEUED"
RUNNING = "RUNNING"
DONE = "DONE"
REJECTED_INVALID = "REJECTED_INVALID"
REJECTED_QUEUE = "REJECTED_QUEUE"
REJECTED_STREAM = "REJECTED_STREAM"
CANCELLED = "CANCELLED"
FAILED = "FAILED"
# iteration counts the served kernel reports for failed requests
SERVE_NON_FINITE = -2
SERVE_BAD_START = -3
FINAL = frozenset({DONE, REJECTED_INVALID, REJECTED_QUEUE, REJECTED_STREAM, CANCELLED, FAILED})


@dataclass
class Ticket:
    slot: int
    ticket: int
    u: Any
    hist: Any
    counters: Any
    n_iters: int | None = None
    passes: int | None = None
    # device buffers the running kernel reads (frame, plan): held until the
    # request is reported done so the caller cannot free them under the kernel
    _keep: tuple = ()


def _to(a, dev):
    """Contiguous f32 tensor on ``dev`` (copies only when needed)."""
    t = _lib.torch()
    if isinstance(a, t.Tensor):
        x = a.to(device=dev, dtype=t.float32)
    else:
        x = t.as_tensor(np.asarray(a, dtype=np.float32), device=dev)
    return x.contiguous()


def prepare_inputs(n_nodes: int, def_image, u, dev):
    """The served kernel reads raw device pointers, so every input is made a
    contiguous f32 tensor on the session's device first (host arrays, other
    dtypes and strided views are copied) and its shape is checked: the image
    must be 2-D and u must hold (n_nodes, 2) values.  Returns (image, u)."""
    g = _to(def_image, dev)
    if g.dim() != 2:
        raise ValueError("def_image must be a 2-D array (H, W)")
    ud = _to(u, dev)
    if ud.numel() != 2 * int(n_nodes) or tuple(ud.shape) not in ((int(n_nodes), 2), (2 * int(n_nodes),)):
        raise ValueError(f"u must have shape ({int(n_nodes)}, 2), got {tuple(ud.shape)}")
    return g, ud.reshape(int(n_nodes), 2)


class ServeSession:
    """One served session (``d2ic_serve_create``): up to ``slots`` requests in flight."""

    def __init__(self, slots: int = 64, idle_ms: float = 5.0):
        _lib.require_gpu()
        h = ctypes.c_void_p(0)
        _lib.call("d2ic_serve_create", int(slots), float(idle_ms), ctypes.byref(h))
        self.handle = h
        self.slots = int(slots)
        self._lock = threading.Lock()

    def submit(self, plan: CGPlan, def_image, u, reg_strength: float, max_iter: int, tol: float,
               save_history: bool = False, u0=None, counters=None) -> Ticket | None:
        """Queue one solve; the returned ticket's ``u`` receives the solution
        (``u`` itself when it is already a contiguous f32 device tensor of shape
        (Nn, 2)).  The initial guess is ``u0`` when given (only read), else
        ``u``'s contents.  ``def_image`` is converted to a contiguous f32 device
        tensor.  ``counters`` (2 int32 on the device, reusable) receives
        n_iters and the pass count.  Returns None when every slot is busy.

        Device inputs that are already contiguous f32 tensors start no GPU
        work here: while the served kernel holds the GPU, other kernels may
        wait for it to leave, so the request path avoids them."""
        t = _lib.torch()
        g, ud = prepare_inputs(plan.n_nodes, def_image, u, _lib.device())
        u0d = None
        if u0 is not None:
            _, u0d = prepare_inputs(plan.n_nodes, g, u0, _lib.device())
        hist = None
        if save_history:
            hist = t.full((max(int(max_iter), 0), 3), float("nan"), dtype=t.float32, device=_lib.device())
        if counters is None:
            counters = t.empty((2,), dtype=t.int32, device=_lib.device())
        t.cuda.current_stream().synchronize()  # inputs must be on the device before the kernel reads them
        slot = ctypes.c_int32(-1)
        ticket = ctypes.c_int64(0)
        with self._lock:
            _lib.call("d2ic_serve_submit", self.handle, plan.handle, _lib.ptr(g), int(g.shape[0]),
                      int(g.shape[1]), _lib.ptr(ud), _lib.ptr(u0d), float(reg_strength), int(max_iter), float(tol),
                      _lib.ptr(hist), _lib.ptr(counters), ctypes.c_void_p(counters.data_ptr() + 4),
                      ctypes.byref(slot), ctypes.byref(ticket))
        if slot.value < 0:
            return None
        return Ticket(slot.value, ticket.value, ud, hist, counters, _keep=(g, plan, u0d))

    def poll(self, tk: Ticket) -> bool:
        done = ctypes.c_int32(0)
        it = ctypes.c_int32(0)
        ps = ctypes.c_int32(0)
        with self._lock:
            _lib.call("d2ic_serve_poll", self.handle, tk.slot, tk.ticket, ctypes.byref(done), ctypes.byref(it),
                      ctypes.byref(ps))
        if done.value:
            tk.n_iters, tk.passes = int(it.value), int(ps.value)
            tk._keep = ()
            return True
        return False

    def wait(self, tk: Ticket, timeout_s: float = 60.0) -> Ticket:
        t0 = time.perf_counter()
        while not self.poll(tk):
            if time.perf_counter() - t0 > timeout_s:
                raise RuntimeError("served request did not finish in time")
            time.sleep(20e-6)
        return tk

    def close(self):
        """Stop the served kernel (after its running requests) and free the session."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            _lib.load().d2ic_serve_destroy(h)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


@dataclass
class StreamState:
    key: Any
    plan: Any
    nodes: int
    image_shape: tuple
    reg_strength: float
    max_iter: int
    tol: float
    u_prev: Any = None
    counters: Any = None  # n_iters and pass count of the request in flight (device)
    bytes_resident: int = 0
    queue: deque = field(default_factory=deque)
    inflight: Any = None
    closing: bool = False


@dataclass
class Request:
    key: Any
    frame: Any
    t_arrival: float
    status: str = QUEUED
    reason: str = ""
    u0: Any = None
    t_submit: float | None = None
    t_done: float | None = None
    u: Any = None
    n_iters: int | None = None
    passes: int | None = None
    _cancel: bool = False


def _all_finite(x) -> bool:
    t = _lib.torch()
    if isinstance(x, t.Tensor):
        return bool(t.isfinite(x).all().item())
    return bool(np.isfinite(np.asarray(x)).all())


def _on_device(x) -> bool:
    t = _lib.torch()
    return isinstance(x, t.Tensor) and x.device.type != "cpu"


class StreamServer:
    """Frame-pair solve server for many image streams (continuous batching).

    ``add_stream`` builds the stream's plan (pixel layout, f sampled once) and
    keeps it, the reference samples and the warm start resident in HBM; streams
    are admitted while their device bytes fit ``hbm_budget_bytes``.  ``submit``
    validates a frame and enqueues it (at most ``max_queue`` waiting frames per
    stream); ``pump`` submits the head frame of every idle stream and collects
    finished ones (frame k+1 of a stream starts from frame k's solution).
    Latency = completion time - arrival time.  Every request ends in one of
    the final status codes listed in the module docstring.

    ``session`` and ``plan_factory`` replace the served HIP session and the
    plan constructor (both default to the real ones); they exist so that the
    host logic can be exercised without a GPU."""

    def __init__(self, slots: int = 64, hbm_budget_bytes: float = 200e9, idle_ms: float = 5.0,
                 max_queue: int = 64, check_finite: bool = True, session: Any = None,
                 plan_factory: Callable[[Any], Any] | None = None, device: Any = None):
        self.session = session if session is not None else ServeSession(slots=slots, idle_ms=idle_ms)
        self.device = device if device is not None else _lib.device()
        self._plan_factory = plan_factory if plan_factory is not None else CGPlan
        self.budget = float(hbm_budget_bytes)
        self.max_queue = int(max_queue)
        self.check_finite = bool(check_finite)
        self.used = 0
        self.streams: dict[Any, StreamState] = {}
        self.done: list[Request] = []      # requests that ended DONE
        self.unsolved: list[Request] = []  # requests that ended rejected, cancelled or failed

    # ------------------------------------------------------------------ streams
    def add_stream(self, key, pixel_assets, ref_image, *, reg_strength=0.1, max_iter=200, tol=1e-3,
                   interpolation="cubic", u0=None) -> bool:
        """Admit a stream if its resident state fits the HBM budget (returns
        False otherwise, and nothing stays allocated)."""
        if key in self.streams:
            raise ValueError(f"stream {key!r} already exists")
        interp = _normalize_interpolation(interpolation)
        t = _lib.torch()
        plan = self._plan_factory(pixel_assets)
        need = int(plan.info()["device_bytes"]) + 8 * int(plan.n_nodes)
        if self.used + need > self.budget:
            plan.close()
            return False
        ref = plan.set_reference(ref_image, interp)
        shape = tuple(int(v) for v in np.shape(ref if ref is not None else ref_image))
        st = StreamState(key, plan, int(plan.n_nodes), shape, float(reg_strength), int(max_iter), float(tol))
        if u0 is not None:
            st.u_prev = _to(u0, self.device).reshape(plan.n_nodes, 2).clone()
        else:
            st.u_prev = t.zeros((plan.n_nodes, 2), dtype=t.float32, device=self.device)
        st.counters = t.empty((2,), dtype=t.int32, device=self.device)
        st.bytes_resident = need
        self.used += need
        self.streams[key] = st
        return True

    def remove_stream(self, key, *, wait: bool = True, timeout_s: float = 60.0):
        """Drop a stream: its queued frames are cancelled; a frame in flight is
        waited for (``wait=True``, the default) or the call raises, since the
        served kernel still reads the stream's plan.  Frees the plan after."""
        st = self.streams.get(key)
        if st is None:
            raise KeyError(key)
        st.closing = True
        while st.queue:
            self._finish(st.queue.popleft(), CANCELLED, "stream removed")
        if st.inflight is not None:
            if not wait:
                st.closing = False
                raise RuntimeError(f"stream {key!r} has a request in flight")
            t0 = time.perf_counter()
            while st.inflight is not None:
                self._collect(st)
                if st.inflight is not None:
                    if time.perf_counter() - t0 > timeout_s:
                        raise RuntimeError(f"stream {key!r}: in-flight request did not finish in time")
                    time.sleep(20e-6)
        del self.streams[key]
        self.used -= st.bytes_resident
        st.plan.close()

    # ----------------------------------------------------------------- requests
    def _finish(self, req: Request, status: str, reason: str = "") -> Request:
        req.status = status
        req.reason = reason
        if req.t_done is None:
            req.t_done = time.perf_counter()
        (self.done if status == DONE else self.unsolved).append(req)
        return req

    def submit(self, key, frame, t_arrival: float | None = None) -> Request:
        """Validate and enqueue one frame of stream ``key``; the returned
        request's ``status`` tells whether it was queued or rejected."""
        req = Request(key, frame, time.perf_counter() if t_arrival is None else t_arrival)
        st = self.streams.get(key)
        if st is None or st.closing:
            return self._finish(req, REJECTED_STREAM, f"unknown stream {key!r}")
        shape = tuple(int(v) for v in np.shape(frame))
        if len(shape) != 2 or shape != st.image_shape:
            return self._finish(req, REJECTED_INVALID,
                                f"frame shape {shape} differs from the reference image {st.image_shape}")
        # host frames are checked here; a device frame is checked by the served
        # kernel's first pass (a check here would start GPU work, which can wait
        # for the served kernel to leave the GPU)
        if self.check_finite and not _on_device(frame) and not _all_finite(frame):
            return self._finish(req, REJECTED_INVALID, "frame holds NaN or Inf values")
        if len(st.queue) >= self.max_queue:
            return self._finish(req, REJECTED_QUEUE, f"stream queue full ({self.max_queue} frames)")
        st.queue.append(req)
        return req

    def cancel(self, req: Request) -> bool:
        """Cancel a request.  A queued one leaves its queue now; a running one
        finishes on the device but its result is discarded (the stream keeps
        its previous warm start).  Returns False if it had already ended."""
        if req.status in FINAL:
            return False
        st = self.streams.get(req.key)
        if req.status == QUEUED and st is not None and req in st.queue:
            st.queue.remove(req)
            self._finish(req, CANCELLED, "cancelled while queued")
            return True
        req._cancel = True
        return True

    def _collect(self, st: StreamState) -> int:
        if st.inflight is None:
            return 0
        req, tk = st.inflight
        if not self.session.poll(tk):
            return 0
        st.inflight = None
        req.t_done = time.perf_counter()
        req.n_iters, req.passes = tk.n_iters, tk.passes
        if req._cancel:
            self._finish(req, CANCELLED, "cancelled while running")
        elif tk.n_iters == SERVE_BAD_START:
            self._finish(req, REJECTED_INVALID,
                         "the frame or the warm start holds NaN or Inf values where the solve samples it")
        elif tk.n_iters == SERVE_NON_FINITE:
            self._finish(req, FAILED, "non-finite displacements (the final objective or gradient is not finite)")
        elif tk.n_iters is None or tk.n_iters < 0:
            self._finish(req, FAILED, "the solve did not finish (a grid barrier timed out)")
        else:
            req.u = tk.u
            st.u_prev = tk.u
            self._finish(req, DONE)
        return 1

    def pump(self) -> int:
        """Submit waiting head frames, collect finished ones; returns the number
        of requests that ended in this call."""
        n = 0
        for st in list(self.streams.values()):
            n += self._collect(st)
            if st.inflight is None and st.queue and not st.closing:
                req = st.queue[0]
                # the solve reads the warm start from the previous result and
                # writes a fresh buffer: no device work on the request path
                u = _lib.torch().empty_like(st.u_prev)
                tk = self.session.submit(st.plan, req.frame, u, st.reg_strength, st.max_iter, st.tol,
                                         u0=st.u_prev, counters=st.counters)
                if tk is not None:
                    st.queue.popleft()
                    req.status = RUNNING
                    req.t_submit = time.perf_counter()
                    st.inflight = (req, tk)
        return n

    def busy(self) -> bool:
        return any(st.inflight is not None or st.queue for st in self.streams.values())

    def drain(self, timeout_s: float = 120.0):
        t0 = time.perf_counter()
        while self.busy():
            self.pump()
            if time.perf_counter() - t0 > timeout_s:
                raise RuntimeError("StreamServer.drain timed out")
            time.sleep(20e-6)

    def latencies(self) -> np.ndarray:
        return np.array([r.t_done - r.t_arrival for r in self.done if r.t_done is not None])

    def close(self, timeout_s: float = 60.0):
        """Cancel queued frames, wait for the requests the served kernel is
        running (it reads their plans), stop the kernel, then free every
        stream's plan."""
        for st in self.streams.values():
            st.closing = True
            while st.queue:
                self._finish(st.queue.popleft(), CANCELLED, "server closed")
        t0 = time.perf_counter()
        while any(st.inflight is not None for st in self.streams.values()):
            for st in self.streams.values():
                self._collect(st)
            if time.perf_counter() - t0 > timeout_s:
                break
            time.sleep(20e-6)
        self.session.close()  # the served kernel has left when this returns
        for st in self.streams.values():
            if st.inflight is not None:
                req, _ = st.inflight
                st.inflight = None
                self._finish(req, FAILED, "server closed before the request finished")
            st.plan.close()
        self.streams.clear()
        self.used = 0
