"""Serving many device-resident maps: Python side of the request loop in
``csrc/ikd_serve.hip`` (``ikd_server_*`` in ``include/ikd_tree_amd.h``).

A request is one scan for one map: kNN of every scan point against the map,
then Add_Points of the scan, then an optional box trim.  The native server
runs ``slots`` worker threads that admit requests as slots and HBM budget free
up; requests for one map run in submission order, requests for different maps
overlap on the device.

``simulate`` drives the server with Poisson arrivals from many simulated
robots (each with its own map and LiDAR-like scene) in real time and reports
p50/p99 of completion minus arrival, next to throughput.
"""
from __future__ import annotations

import ctypes
import importlib
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_pkg = importlib.import_module("ikd-tree_amd")


class Request(ctypes.Structure):
    _fields_ = [
        ("map", ctypes.c_int32), ("n", ctypes.c_int32), ("xyz", ctypes.c_void_p),
        ("k", ctypes.c_int32), ("max_dist", ctypes.c_double), ("downsample", ctypes.c_int32),
        ("n_trim", ctypes.c_int32), ("trim_boxes", ctypes.c_void_p),
        ("out_xyz", ctypes.c_void_p), ("out_d2", ctypes.c_void_p), ("out_ids", ctypes.c_void_p),
        ("out_counts", ctypes.c_void_p), ("t_arrival_ns", ctypes.c_int64), ("tag", ctypes.c_int64),
        ("t_start_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64),
        ("added", ctypes.c_int32), ("trimmed", ctypes.c_int32), ("status", ctypes.c_int32),
        ("t_insert_done_ns", ctypes.c_int64), ("delta", _pkg.Counters), ("flags", ctypes.c_int32),
    ]


REQ_PINNED_RESULTS = 1  # IKD_REQ_PINNED_RESULTS


def _lib():
    lib = _pkg.load_library()
    if not getattr(lib, "_serve_sigs", False):
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        lib.ikd_now_ns.argtypes = []
        lib.ikd_now_ns.restype = i64
        for name, args in {
            "ikd_server_create": [i32, i64, ctypes.POINTER(vp)],
            "ikd_server_add_map": [vp, vp, ctypes.POINTER(i32)],
            "ikd_server_submit": [vp, ctypes.POINTER(Request)],
            "ikd_server_wait": [vp, ctypes.POINTER(ctypes.POINTER(Request)), i32, i32, ctypes.POINTER(i32)],
            "ikd_server_pending": [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)],
            "ikd_server_destroy": [vp],
            "ikd_server_set_maintenance": [vp, i64],
            "ikd_server_stats": [vp, ctypes.POINTER(i64)],
            "ikd_server_memory": [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)],
            "ikd_host_alloc": [i64, ctypes.POINTER(vp)],
            "ikd_host_free": [vp],
        }.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = i32
        lib._serve_sigs = True
    return lib


def now_ns() -> int:
    return int(_lib().ikd_now_ns())


class _PinnedBlock:
    """Page-locked host memory from ikd_host_alloc, freed with the object."""

    def __init__(self, nbytes: int):
        self._lib = _lib()
        p = ctypes.c_void_p()
        rc = self._lib.ikd_host_alloc(int(nbytes), ctypes.byref(p))
        if rc != 0:
            raise _pkg.IkdError(rc, "page-locked allocation failed")
        self.ptr = p.value

    def array(self, offset: int, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        raw = (ctypes.c_byte * max(n, 1)).from_address(self.ptr + offset)
        raw.block = self  # the arrays keep the memory alive
        return np.frombuffer(raw, dt, count=int(np.prod(shape))).reshape(shape)

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.ikd_host_free(self.ptr)
            self.ptr = None


class ResultBuffers(tuple):
    """(out_xyz, out_d2, out_ids, out_counts) for requests of up to n_max
    queries at k neighbours; .pinned: page-locked (ikd_host_alloc)."""
    pinned = False


def result_buffers(n_max: int, k: int, pinned: bool = False) -> ResultBuffers:
    """Host result arrays for requests of up to n_max queries at k neighbours,
    written once so that their pages are resident: a client that reuses them
    (submit(..., outs=...)) keeps page faults out of the workers' result
    copies.  pinned=True puts them in page-locked memory, which the server
    fills by DMA with no host copy of its own."""
    shapes = (((n_max, k, 3), np.float32), ((n_max, k), np.float32), ((n_max, k), np.uint32),
              ((n_max,), np.int32))
    if pinned:
        sizes = [int(np.prod(sh)) * np.dtype(dt).itemsize for sh, dt in shapes]
        offs = [0]
        for sz in sizes[:-1]:
            offs.append(offs[-1] + (sz + 255) // 256 * 256)
        block = _PinnedBlock(offs[-1] + sizes[-1] + 256)
        outs = ResultBuffers(block.array(o, sh, dt) for o, (sh, dt) in zip(offs, shapes))
        outs.pinned = True
    else:
        outs = ResultBuffers(np.empty(sh, dt) for sh, dt in shapes)
    for a in outs:
        a.fill(0)
    return outs


class Ticket:
    """One submitted request: keeps its buffers alive and exposes the results.
    With outs (arrays from result_buffers, at least n rows and the request's
    k) the results are written there instead of into new arrays; out_* are
    then views of outs, valid until the caller hands outs to another request."""

    def __init__(self, map_index, xyz, k, max_dist, downsample, trim, tag, outs=None):
        self.xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = self.xyz.shape[0]
        self.k = int(k)
        kk = max(self.k, 0)
        self.outs = outs
        if outs is not None:
            ox, od, oi, oc = outs
            if (ox.shape[0] < n or ox.shape[1:] != (kk, 3) or od.shape[1:] != (kk,) or oi.shape[1:] != (kk,)
                    or min(od.shape[0], oi.shape[0], oc.shape[0]) < n or ox.dtype != np.float32
                    or od.dtype != np.float32 or oi.dtype != np.uint32 or oc.dtype != np.int32
                    or not all(a.flags.c_contiguous for a in outs)):
                raise ValueError("result buffers do not fit the request")
            self.out_xyz, self.out_d2, self.out_ids, self.out_counts = ox[:n], od[:n], oi[:n], oc[:n]
            pinned = bool(getattr(outs, "pinned", False))
        else:
            self.out_xyz = np.zeros((n, kk, 3), np.float32)
            self.out_d2 = np.zeros((n, kk), np.float32)
            self.out_ids = np.zeros((n, kk), np.uint32)
            self.out_counts = np.zeros(n, np.int32)
            pinned = False
        self.trim = None if trim is None else np.ascontiguousarray(trim, np.float32).reshape(-1, 6)
        r = Request()
        r.map = map_index
        r.n = n
        r.xyz = self.xyz.ctypes.data
        r.k = self.k
        r.max_dist = float(max_dist)
        r.downsample = int(downsample)
        r.n_trim = 0 if self.trim is None else self.trim.shape[0]
        r.trim_boxes = None if self.trim is None else self.trim.ctypes.data
        if kk > 0 and n > 0:
            r.out_xyz = self.out_xyz.ctypes.data
            r.out_d2 = self.out_d2.ctypes.data
            r.out_ids = self.out_ids.ctypes.data
            r.out_counts = self.out_counts.ctypes.data
        r.tag = int(tag)
        r.flags = REQ_PINNED_RESULTS if pinned else 0
        self.req = r

    @property
    def latency_ms(self) -> float:
        return (self.req.t_done_ns - self.req.t_arrival_ns) / 1e6

    @property
    def status(self) -> int:
        return int(self.req.status)


class Server:
    def __init__(self, slots: int = 8, hbm_budget_bytes: int = 64 << 30):
        self._lib = _lib()
        h = ctypes.c_void_p()
        rc = self._lib.ikd_server_create(slots, int(hbm_budget_bytes), ctypes.byref(h))
        if rc != 0:
            raise _pkg.IkdError(rc, "server create failed")
        self._h = h
        self._tickets = {}
        self._maps = []

    def set_maintenance(self, max_points: int):
        """Deferred-rebuild budget the workers spend on an idle map (-1: no limit, 0: never)."""
        rc = self._lib.ikd_server_set_maintenance(self._h, int(max_points))
        if rc != 0:
            raise _pkg.IkdError(rc, "set_maintenance failed")

    def maintenance_rebuilds(self) -> int:
        v = ctypes.c_int64(0)
        self._lib.ikd_server_stats(self._h, ctypes.byref(v))
        return int(v.value)

    def memory(self):
        """(device bytes of the registered maps, estimate of the running requests)."""
        res, fl = ctypes.c_int64(0), ctypes.c_int64(0)
        rc = self._lib.ikd_server_memory(self._h, ctypes.byref(res), ctypes.byref(fl))
        if rc != 0:
            raise _pkg.IkdError(rc, "memory failed")
        return int(res.value), int(fl.value)

    @staticmethod
    def request_bytes(n: int, k: int, downsample: int) -> int:
        """The server's device-memory estimate of one request (scan, kNN
        results, insert scratch)."""
        return n * 12 + n * max(k, 0) * 20 + n * 4 + (n * 160 if downsample >= 0 else 0) + (1 << 20)

    def add_map(self, tree, defer_rebuild_points: int | None = None) -> int:
        """Register a map; with defer_rebuild_points its large rebuilds are
        deferred to the idle time between its requests."""
        if defer_rebuild_points is not None:
            tree.set_tuning(_pkg.IKD_TUNE_DEFER_REBUILD, int(defer_rebuild_points))
        m = ctypes.c_int32(0)
        rc = self._lib.ikd_server_add_map(self._h, tree._h, ctypes.byref(m))
        if rc != 0:
            raise _pkg.IkdError(rc, "add_map failed")
        self._maps.append(tree)
        return int(m.value)

    def submit(self, map_index: int, xyz, k: int = 5, max_dist: float = 5.0, downsample: int = 1, trim=None,
               tag: int = 0, arrival_ns: int | None = None, outs=None) -> Ticket:
        t = Ticket(map_index, xyz, k, max_dist, downsample, trim, tag, outs)
        t.req.t_arrival_ns = now_ns() if arrival_ns is None else int(arrival_ns)
        rc = self._lib.ikd_server_submit(self._h, ctypes.byref(t.req))
        if rc != 0:
            raise _pkg.IkdError(rc, "submit rejected")
        self._tickets[ctypes.addressof(t.req)] = t
        return t

    def wait(self, timeout_ms: int = -1, max_out: int = 64):
        arr = (ctypes.POINTER(Request) * max_out)()
        n = ctypes.c_int32(0)
        rc = self._lib.ikd_server_wait(self._h, arr, max_out, int(timeout_ms), ctypes.byref(n))
        if rc != 0:
            raise _pkg.IkdError(rc, "wait failed")
        return [self._tickets.pop(ctypes.addressof(arr[i].contents)) for i in range(n.value)]

    def outstanding(self) -> int:
        return len(self._tickets)

    def drain(self):
        out = []
        while self._tickets:
            out.extend(self.wait(-1))
        return out

    def wait_idle(self, timeout_s: float = 60.0):
        """Waits until no request is pending or running (workers hand a request
        back before they finish the map's maintenance and accounting)."""
        pend, fl = ctypes.c_int64(0), ctypes.c_int64(0)
        end = time.monotonic() + timeout_s
        while True:
            rc = self._lib.ikd_server_pending(self._h, ctypes.byref(pend), ctypes.byref(fl))
            if rc != 0:
                raise _pkg.IkdError(rc, "pending failed")
            if pend.value == 0 and fl.value == 0:
                return
            if time.monotonic() > end:
                raise TimeoutError("server did not go idle")
            time.sleep(0.001)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ikd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate(server: Server, map_ids, scan_fn, rate_hz: float = 10.0, duration_s: float = 5.0, k: int = 5,
             max_dist: float = 5.0, trim_fn=None, trim_every: int = 10, seed: int = 1005,
             pinned_results: bool = True):
    """Real-time Poisson arrivals: robot r's scans arrive at rate_hz (exponential
    gaps) for duration_s seconds; scan_fn(r, j) gives robot r's j-th scan and
    trim_fn(r, j) its trim boxes (every trim_every-th request).  Each robot's
    client reuses its result buffers (page-locked with pinned_results).
    Returns latencies (ms) and throughput."""
    rng = np.random.default_rng(seed)
    events = []
    for r, m in enumerate(map_ids):
        t, j = 0.0, 0
        while True:
            t += rng.exponential(1.0 / rate_hz)
            if t >= duration_s:
                break
            events.append((t, r, m, j))
            j += 1
    events.sort()
    keys = [(r, j) for (_, r, _, j) in events]
    with ThreadPoolExecutor(max_workers=16) as ex:  # scan synthesis releases the GIL
        scans = dict(zip(keys, ex.map(lambda rj: scan_fn(*rj), keys)))
    # each robot's result buffers are reused once its request is back, as a
    # client keeps them (two sets per robot to start; more if requests queue)
    n_max = max(s.shape[0] for s in scans.values())
    free = {r: [result_buffers(n_max, k, pinned_results) for _ in range(2)] for r in range(len(map_ids))}

    def collect(got):
        for d in got:
            free[d.req.tag].append(d.outs)
            d.outs = None
        done.extend(got)

    t0 = now_ns() + 50_000_000
    done = []
    points = 0
    def wait_until(due):
        """Collect finished requests until the clock reaches due, submitting
        late by tens of microseconds at most (a late submission would count as
        queue wait): ms-granular waits while far, short sleeps, then a spin."""
        while True:
            left = (due - now_ns()) / 1e6
            if left <= 0:
                return
            if left > 1.5:
                t_call = now_ns()
                got = server.wait(timeout_ms=int(left - 0.5))
                collect(got)
                if not got and now_ns() - t_call < 200_000:  # idle server: wait() returned at once
                    time.sleep(min(left - 0.5, 1.0) / 1e3)
            elif left > 0.3:
                collect(server.wait(timeout_ms=0))
                time.sleep((left - 0.25) / 1e3)
            else:
                collect(server.wait(timeout_ms=0))

    for (t, r, m, j) in events:
        due = t0 + int(t * 1e9)
        wait_until(due)
        trim = trim_fn(r, j) if (trim_fn is not None and j % trim_every == trim_every - 1) else None
        s = scans[(r, j)]
        points += s.shape[0]
        outs = free[r].pop() if free[r] else result_buffers(n_max, k, pinned_results)
        server.submit(m, s, k, max_dist, 1, trim, tag=r, arrival_ns=due, outs=outs)
    collect(server.drain())
    t1 = max(d.req.t_done_ns for d in done)
    lat = np.array([d.latency_ms for d in done])
    wait = np.array([(d.req.t_start_ns - d.req.t_arrival_ns) / 1e6 for d in done])
    run = np.array([(d.req.t_done_ns - d.req.t_start_ns) / 1e6 for d in done])
    bad = [d for d in done if d.status != 0]
    slow = sorted(done, key=lambda d: -d.latency_ms)[:8]

    def pct(a):
        return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)), "max": float(a.max())}

    # latency percentiles per 5 s of arrivals: how the tail moves over the run
    arr_s = np.array([(d.req.t_arrival_ns - t0) / 1e9 for d in done])
    by_window = []
    for w0 in np.arange(0.0, duration_s, 5.0):
        sel = lat[(arr_s >= w0) & (arr_s < w0 + 5.0)]
        if sel.size:
            by_window.append({"from_s": float(w0), "requests": int(sel.size), "p50": float(np.percentile(sel, 50)),
                              "p99": float(np.percentile(sel, 99))})

    return {
        "requests": len(done),
        "failed": len(bad),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max()), "mean": float(lat.mean())},
        "pinned_results": bool(pinned_results),
        "queue_wait_ms": pct(wait),
        "run_ms": pct(run),
        "latency_ms_by_window": by_window,
        "slowest": [{"robot": int(d.req.tag), "points": int(d.req.n), "trim": int(d.req.n_trim > 0),
                     "wait_ms": (d.req.t_start_ns - d.req.t_arrival_ns) / 1e6,
                     "run_ms": (d.req.t_done_ns - d.req.t_start_ns) / 1e6,
                     "insert_done_ms": (d.req.t_insert_done_ns - d.req.t_start_ns) / 1e6,
                     "events": d.req.delta.as_dict()} for d in slow],
        "events": {
            "requests_with_compaction": sum(1 for d in done if d.req.delta.compactions > 0),
            "requests_with_pool_grow": sum(1 for d in done if d.req.delta.pool_grows > 0),
            "requests_with_scratch_grow": sum(1 for d in done if d.req.delta.scratch_grows > 0),
            "rebuilt_points_per_request": float(np.mean([d.req.delta.rebuilt_points for d in done])),
            "host_waits_per_request": float(np.mean([d.req.delta.host_waits for d in done])),
        },
        "scan_points_per_s": points / ((t1 - t0) / 1e9),
        "requests_per_s": len(done) / ((t1 - t0) / 1e9),
        "wall_s": (t1 - t0) / 1e9,
    }
