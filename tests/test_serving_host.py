"""StreamServer host logic on the CPU: admission against the HBM budget,
request status codes, bounded queues, NaN-frame rejection, cancellation,
failed solves, warm-start chaining and the teardown order.  A recording
stand-in replaces the served HIP session and the device plan, so no GPU is
touched; the GPU tests check the served kernel itself."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from d2ic_amd import serving as S


class FakePlan:
    def __init__(self, spec):
        self.n_nodes = int(spec["n_nodes"])
        self.bytes = int(spec.get("bytes", 1000))
        self.closed = False
        self.events = spec.setdefault("events", [])

    def info(self):
        return {"device_bytes": self.bytes}

    def set_reference(self, ref, interp):
        return torch.as_tensor(np.asarray(ref, np.float32))

    def close(self):
        self.closed = True
        self.events.append("plan.close")


class FakeTicket:
    def __init__(self, u, polls_left, outcome):
        self.u, self.polls_left, self.outcome = u, polls_left, outcome
        self.n_iters = self.passes = None


class FakeSession:
    """Stands in for ServeSession: ``slots`` requests at a time; a request is
    done after ``latency`` polls and then holds u0 + 1 (or the outcome that
    ``plan_outcome`` names for the frame's first pixel value)."""

    def __init__(self, slots=4, latency=2, events=None):
        self.slots, self.latency = slots, latency
        self.running = []
        self.submitted = []
        self.events = events if events is not None else []

    def submit(self, plan, frame, u, reg, max_iter, tol, u0=None, counters=None):
        if len(self.running) >= self.slots:
            return None
        first = float(np.asarray(frame).reshape(-1)[0])
        outcome = {-1.0: "timeout", -2.0: "nan", -3.0: "bad_start"}.get(first, "ok")
        start = (u0 if u0 is not None else u).clone()
        tk = FakeTicket(start, self.latency, outcome)
        self.running.append(tk)
        self.submitted.append((plan, frame, start))
        return tk

    def poll(self, tk):
        tk.polls_left -= 1
        if tk.polls_left > 0:
            return False
        self.running.remove(tk)
        # the served kernel's result codes: -1 timeout, -2 non-finite result,
        # -3 non-finite objective at the initial guess
        codes = {"timeout": -1, "nan": -2, "bad_start": -3}
        if tk.outcome in codes:
            tk.n_iters, tk.passes = codes[tk.outcome], 1
        else:
            tk.u = tk.u + 1.0
            tk.n_iters, tk.passes = 3, 4
        return True

    def close(self):
        self.events.append("session.close")


def _server(slots=4, budget=1e9, max_queue=4, latency=2):
    events = []
    sess = FakeSession(slots=slots, latency=latency, events=events)
    srv = S.StreamServer(hbm_budget_bytes=budget, max_queue=max_queue, session=sess, plan_factory=FakePlan,
                         device="cpu")
    return srv, sess, events


def _frame(v=0.0, shape=(8, 8)):
    f = torch.full(shape, 5.0)
    f.view(-1)[0] = v
    return f


REF = np.zeros((8, 8), np.float32)


def test_budget_admission_rejects_without_allocating():
    srv, _, _ = _server(budget=2 * 1000 + 2 * 8 * 3 + 10)
    specs = [dict(n_nodes=3) for _ in range(3)]
    assert srv.add_stream("a", specs[0], REF)
    assert srv.add_stream("b", specs[1], REF)
    used = srv.used
    assert not srv.add_stream("c", specs[2], REF)  # over the budget
    assert srv.used == used and "c" not in srv.streams
    assert specs[2]["events"] == ["plan.close"]  # its plan was released at once
    with pytest.raises(ValueError):
        srv.add_stream("a", dict(n_nodes=3), REF)


def test_invalid_frames_are_rejected_before_submission():
    srv, sess, _ = _server()
    assert srv.add_stream("a", dict(n_nodes=3), REF)
    bad_shape = srv.submit("a", torch.zeros(8, 9))
    three_d = srv.submit("a", torch.zeros(2, 8, 8))
    nan = srv.submit("a", _frame(float("nan")))
    inf = srv.submit("a", _frame(float("inf")))
    unknown = srv.submit("zz", _frame())
    srv.drain()
    assert [r.status for r in (bad_shape, three_d, nan, inf)] == [S.REJECTED_INVALID] * 4
    assert unknown.status == S.REJECTED_STREAM
    assert sess.submitted == []  # none reached the kernel
    assert all(r.t_done is not None for r in srv.unsolved) and len(srv.unsolved) == 5
    # the stream is still usable
    ok = srv.submit("a", _frame())
    srv.drain()
    assert ok.status == S.DONE


def test_queue_is_bounded():
    srv, _, _ = _server(max_queue=2, latency=50)
    assert srv.add_stream("a", dict(n_nodes=2), REF)
    reqs = [srv.submit("a", _frame()) for _ in range(3)]
    assert [r.status for r in reqs] == [S.QUEUED, S.QUEUED, S.REJECTED_QUEUE]
    srv.pump()  # head goes to the kernel: one queue place frees up
    assert reqs[0].status == S.RUNNING
    assert srv.submit("a", _frame()).status == S.QUEUED
    srv.drain()
    assert [r.status for r in reqs[:2]] == [S.DONE, S.DONE]


def test_warm_start_chains_frames_of_a_stream():
    srv, sess, _ = _server()
    u0 = np.arange(6, dtype=np.float32).reshape(3, 2)
    assert srv.add_stream("a", dict(n_nodes=3), REF, u0=u0)
    reqs = [srv.submit("a", _frame()) for _ in range(3)]
    srv.drain()
    assert [r.status for r in reqs] == [S.DONE] * 3
    for k, (_, _, u_in) in enumerate(sess.submitted):
        np.testing.assert_array_equal(u_in.numpy(), u0 + k)  # frame k starts from frame k-1's result
    np.testing.assert_array_equal(reqs[-1].u.numpy(), u0 + 3)
    assert all(r.n_iters == 3 and r.passes == 4 for r in reqs)


def test_failed_solves_keep_the_previous_warm_start():
    srv, sess, _ = _server(max_queue=8)
    assert srv.add_stream("a", dict(n_nodes=2), REF)
    r1 = srv.submit("a", _frame())
    r2 = srv.submit("a", _frame(-1.0))  # the stand-in reports a barrier time-out
    r3 = srv.submit("a", _frame(-2.0))  # the stand-in reports a non-finite result
    r5 = srv.submit("a", _frame(-3.0))  # ... a non-finite objective at the initial guess
    r4 = srv.submit("a", _frame())
    srv.drain()
    assert [r.status for r in (r1, r2, r3, r5, r4)] == [S.DONE, S.FAILED, S.FAILED, S.REJECTED_INVALID, S.DONE]
    assert "time" in r2.reason and "non-finite" in r3.reason and "NaN or Inf" in r5.reason
    assert all(r.u is None for r in (r2, r3, r5))
    # r2..r4 all started from r1's result: failures did not poison the stream
    for _, _, u_in in sess.submitted[1:]:
        np.testing.assert_array_equal(u_in.numpy(), np.ones((2, 2), np.float32))
    np.testing.assert_array_equal(r4.u.numpy(), 2 * np.ones((2, 2), np.float32))


def test_cancel_queued_and_running():
    srv, sess, _ = _server(latency=3)
    assert srv.add_stream("a", dict(n_nodes=2), REF)
    r1, r2, r3 = (srv.submit("a", _frame()) for _ in range(3))
    srv.pump()
    assert r1.status == S.RUNNING
    assert srv.cancel(r2)  # queued: leaves now
    assert r2.status == S.CANCELLED and len(sess.submitted) == 1
    assert srv.cancel(r1)  # running: finishes, result discarded
    srv.drain()
    assert r1.status == S.CANCELLED and r1.u is None
    assert r3.status == S.DONE
    # r3 started from the initial zeros, not from the cancelled r1's result
    np.testing.assert_array_equal(sess.submitted[1][2].numpy(), np.zeros((2, 2), np.float32))
    assert not srv.cancel(r3)  # already ended


def test_remove_stream_waits_for_the_request_in_flight():
    srv, sess, _ = _server(latency=4)
    spec = dict(n_nodes=2)
    assert srv.add_stream("a", spec, REF)
    r1, r2 = srv.submit("a", _frame()), srv.submit("a", _frame())
    srv.pump()
    with pytest.raises(RuntimeError):
        srv.remove_stream("a", wait=False)
    assert "a" in srv.streams and spec["events"] == []
    srv.remove_stream("a")
    assert r1.status == S.DONE and r2.status == S.CANCELLED
    assert spec["events"] == ["plan.close"] and srv.used == 0 and sess.running == []
    assert srv.submit("a", _frame()).status == S.REJECTED_STREAM


def test_close_stops_the_kernel_before_freeing_plans():
    srv, sess, events = _server(latency=3)
    specs = [dict(n_nodes=2, events=events), dict(n_nodes=2, events=events)]
    for k, sp in enumerate(specs):
        assert srv.add_stream(k, sp, REF)
    a = srv.submit(0, _frame())
    b = srv.submit(0, _frame())
    c = srv.submit(1, _frame())
    srv.pump()
    srv.close()
    assert events[0] == "session.close" and events[1:] == ["plan.close", "plan.close"]
    assert a.status == S.DONE and c.status == S.DONE  # in flight: waited for
    assert b.status == S.CANCELLED  # queued: cancelled
    assert srv.streams == {} and srv.used == 0


def test_prepare_inputs_converts_or_rejects():
    base = np.arange(64, dtype=np.float64).reshape(8, 8)
    view = torch.as_tensor(base).t()  # float64, non-contiguous
    g, u = S.prepare_inputs(3, view, np.zeros(6, np.float64), "cpu")
    assert g.dtype == torch.float32 and g.is_contiguous() and tuple(g.shape) == (8, 8)
    np.testing.assert_array_equal(g.numpy(), base.T.astype(np.float32))
    assert u.dtype == torch.float32 and tuple(u.shape) == (3, 2) and u.is_contiguous()
    g2, _ = S.prepare_inputs(3, base.astype(np.float32), torch.zeros(3, 2), "cpu")  # numpy frame
    assert isinstance(g2, torch.Tensor)
    with pytest.raises(ValueError):
        S.prepare_inputs(3, np.zeros((2, 8, 8)), np.zeros((3, 2)), "cpu")
    with pytest.raises(ValueError):
        S.prepare_inputs(3, base, np.zeros((4, 2)), "cpu")
