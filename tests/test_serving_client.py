"""The simulated serving client (ikd-tree_amd/serving.py, simulate) against a
stand-in server on the CPU: it must submit each scan at its arrival time, not
up to a millisecond later (a late submission is counted as the server's queue
wait), and reuse each robot's result buffers once a request is back."""
import importlib
import time

import numpy as np

serving = importlib.import_module("ikd-tree_amd.serving")


class _FakeServer:
    """Finishes every request 3 ms after it is submitted, so that requests are
    usually in flight when the client waits for the next arrival."""

    def __init__(self):
        self.pending = []
        self.late_ns = []
        self.outs_seen = []

    def submit(self, m, xyz, k, max_dist, downsample, trim, tag=0, arrival_ns=None, outs=None):
        now = serving.now_ns()
        self.late_ns.append(now - arrival_ns)
        self.outs_seen.append(id(outs))
        t = serving.Ticket(m, xyz, k, max_dist, downsample, trim, tag, outs)
        t.req.t_arrival_ns = arrival_ns
        t.req.t_start_ns = now
        t.req.t_done_ns = now + 3_000_000
        self.pending.append(t)
        return t

    def _ready(self):
        now = serving.now_ns()
        out = [t for t in self.pending if t.req.t_done_ns <= now]
        self.pending = [t for t in self.pending if t.req.t_done_ns > now]
        return out

    def wait(self, timeout_ms=-1, max_out=64):
        end = time.monotonic() + max(timeout_ms, 0) / 1e3
        while True:
            got = self._ready()
            if got or not self.pending or time.monotonic() >= end:
                return got
            time.sleep(0.0001)

    def drain(self):
        out = []
        while self.pending:
            out.extend(self.wait(10))
        return out


def test_client_submits_on_time_and_reuses_buffers():
    srv = _FakeServer()
    pts = np.zeros((64, 3), np.float32)
    res = serving.simulate(srv, [0, 1, 2, 3], lambda r, j: pts, rate_hz=50.0, duration_s=2.0, k=5, max_dist=5.0,
                           pinned_results=False)
    late_ms = np.array(srv.late_ns) / 1e6
    assert res["requests"] == len(late_ms) > 300
    # the former whole-millisecond waits put about one scan in five up to 1 ms
    # late (90th percentile 1.2 ms); these pace within 0.03 ms at the 90th
    # percentile even with every host core busy
    assert np.percentile(late_ms, 90) < 0.3, np.percentile(late_ms, [50, 90, 99])
    assert np.mean(late_ms > 0.5) < 0.12, np.percentile(late_ms, [50, 90, 99])
    # a few buffer sets per robot serve all its requests
    assert len(set(srv.outs_seen)) <= 4 * 4
