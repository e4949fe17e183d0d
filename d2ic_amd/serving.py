"""Serving: continuous batching of global CG solves in one persistent kernel.

No counterpart in the reference (which solves one frame per call,
``solver_global_cg.py:155-198``).  Requests (frame pairs of many camera
streams, each with its own mesh/ROI) join the running persistent CG kernel at
its next grid barrier and leave when solved, so solves of different sizes
overlap on the device.  Each request runs exactly the computation of
``GlobalCGSolver.solve`` and gives bitwise the same nodal displacements.

* :class:`ServeSession` -- thin wrapper of ``d2ic_serve_*`` (slots, tickets).
* :class:`StreamServer` -- per-stream resident state (plan with f sampled at the
  pixels, warm start), admission of streams against an HBM budget, per-stream
  frame ordering (frame k+1 starts from frame k's solution), latency records.
"""
from __future__ import annotations

import ctypes
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _lib
from .solver_global_cg import CGPlan, _normalize_interpolation


@dataclass
class Ticket:
    slot: int
    ticket: int
    u: Any
    hist: Any
    counters: Any
    n_iters: int | None = None
    passes: int | None = None


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
               save_history: bool = False) -> Ticket | None:
        """Queue one solve; ``u`` (device, (Nn,2) f32) holds u0 and receives the solution.
        Returns None when every slot is busy."""
        t = _lib.torch()
        hist = None
        if save_history:
            hist = t.full((max(int(max_iter), 0), 3), float("nan"), dtype=t.float32, device=_lib.device())
        counters = t.full((2,), -1, dtype=t.int32, device=_lib.device())
        t.cuda.current_stream().synchronize()  # inputs must be on the device before the kernel reads them
        slot = ctypes.c_int32(-1)
        ticket = ctypes.c_int64(0)
        with self._lock:
            _lib.call("d2ic_serve_submit", self.handle, plan.handle, _lib.ptr(def_image), int(def_image.shape[0]),
                      int(def_image.shape[1]), _lib.ptr(u), float(reg_strength), int(max_iter), float(tol),
                      _lib.ptr(hist), _lib.ptr(counters), ctypes.c_void_p(counters.data_ptr() + 4),
                      ctypes.byref(slot), ctypes.byref(ticket))
        if slot.value < 0:
            return None
        return Ticket(slot.value, ticket.value, u, hist, counters)

    def poll(self, tk: Ticket) -> bool:
        done = ctypes.c_int32(0)
        it = ctypes.c_int32(0)
        ps = ctypes.c_int32(0)
        with self._lock:
            _lib.call("d2ic_serve_poll", self.handle, tk.slot, tk.ticket, ctypes.byref(done), ctypes.byref(it),
                      ctypes.byref(ps))
        if done.value:
            tk.n_iters, tk.passes = int(it.value), int(ps.value)
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
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            _lib.load().d2ic_serve_destroy(h)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        import sys

        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


@dataclass
class StreamState:
    key: Any
    plan: CGPlan
    nodes: int
    reg_strength: float
    max_iter: int
    tol: float
    u_prev: Any = None
    bytes_resident: int = 0
    queue: deque = field(default_factory=deque)
    inflight: Any = None


@dataclass
class Request:
    key: Any
    frame: Any
    t_arrival: float
    u0: Any = None
    t_submit: float | None = None
    t_done: float | None = None
    u: Any = None
    n_iters: int | None = None
    passes: int | None = None


class StreamServer:
    """Frame-pair solve server for many image streams (continuous batching).

    ``add_stream`` builds the stream's plan (pixel layout, f sampled once) and
    keeps it, the reference samples and the warm start resident in HBM; streams
    are admitted while their device bytes fit ``hbm_budget_bytes``.  ``submit``
    enqueues a frame; ``pump`` submits the head frame of every idle stream and
    collects finished ones (frame k+1 of a stream starts from frame k's
    solution).  Latency = completion time - arrival time."""

    def __init__(self, slots: int = 64, hbm_budget_bytes: float = 200e9, idle_ms: float = 5.0):
        self.session = ServeSession(slots=slots, idle_ms=idle_ms)
        self.budget = float(hbm_budget_bytes)
        self.used = 0
        self.streams: dict[Any, StreamState] = {}
        self.done: list[Request] = []

    def add_stream(self, key, pixel_assets, ref_image, *, reg_strength=0.1, max_iter=200, tol=1e-3,
                   interpolation="cubic", u0=None) -> bool:
        """Admit a stream if its resident state fits the HBM budget (returns False otherwise)."""
        t = _lib.torch()
        plan = CGPlan(pixel_assets)
        need = plan.info()["device_bytes"] + 8 * plan.n_nodes
        if self.used + need > self.budget:
            plan.close()
            return False
        plan.set_reference(ref_image, _normalize_interpolation(interpolation))
        st = StreamState(key, plan, plan.n_nodes, float(reg_strength), int(max_iter), float(tol))
        st.u_prev = (_lib.to_dev(u0, t.float32, (plan.n_nodes, 2)).clone() if u0 is not None else
                     t.zeros((plan.n_nodes, 2), dtype=t.float32, device=_lib.device()))
        st.bytes_resident = need
        self.used += need
        self.streams[key] = st
        return True

    def remove_stream(self, key):
        st = self.streams.pop(key)
        self.used -= st.bytes_resident
        st.plan.close()

    def submit(self, key, frame, t_arrival: float | None = None) -> Request:
        req = Request(key, frame, time.perf_counter() if t_arrival is None else t_arrival)
        self.streams[key].queue.append(req)
        return req

    def pump(self) -> int:
        """Submit waiting head frames, collect finished ones; returns the number finished."""
        n = 0
        for st in self.streams.values():
            if st.inflight is not None:
                req, tk = st.inflight
                if self.session.poll(tk):
                    req.t_done = time.perf_counter()
                    req.u, req.n_iters, req.passes = tk.u, tk.n_iters, tk.passes
                    st.u_prev = tk.u
                    st.inflight = None
                    self.done.append(req)
                    n += 1
            if st.inflight is None and st.queue:
                req = st.queue[0]
                u = st.u_prev.clone()
                tk = self.session.submit(st.plan, req.frame, u, st.reg_strength, st.max_iter, st.tol)
                if tk is not None:
                    st.queue.popleft()
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

    def close(self):
        for key in list(self.streams):
            self.remove_stream(key)
        self.session.close()
