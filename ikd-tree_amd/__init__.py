"""MI355X ikd-Tree: Python mirror of hku-mars/ikd-Tree's ``KD_TREE<PointType>``.

The tree lives in device memory and every method runs HIP kernels (gfx950)
through the C ABI of ``libikd_hip.so`` (declared in ``include/ikd_tree_amd.h``).
Method names, argument meaning and defaults follow ikd-Tree/ikd_Tree.h:225-248.
There is no CPU fallback: importing this package without the built library
raises ``ImportError``.

Import it by path (the directory name is not a Python identifier)::

    import importlib, sys
    sys.path.insert(0, "/path/to/repo")
    ikd = importlib.import_module("ikd-tree_amd")
    tree = ikd.KD_TREE(0.5, 0.6, 0.5)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libikd_hip.so")
CSRC = os.path.join(_HERE, "csrc")

IKD_OK = 0
IKD_ERR_ARG = -1
IKD_ERR_HIP = -2
IKD_ERR_NOMEM = -3
IKD_ERR_CAPACITY = -4
IKD_ERR_UNSUPPORTED = -5
IKD_ERR_INTERNAL = -6

# every entry point include/ikd_tree_amd.h declares
EXPORTS = (
    "ikd_create", "ikd_destroy", "ikd_set_delete_criterion_param", "ikd_set_balance_criterion_param",
    "ikd_set_downsample_param", "ikd_initialize", "ikd_build", "ikd_nearest_search",
    "ikd_nearest_search_device", "ikd_nearest_search_counters", "ikd_box_search", "ikd_radius_search",
    "ikd_add_points", "ikd_add_points_device", "ikd_delete_points", "ikd_delete_point_boxes",
    "ikd_add_point_boxes", "ikd_size",
    "ikd_validnum", "ikd_tree_range", "ikd_root_alpha", "ikd_flatten", "ikd_synchronize", "ikd_stream",
    "ikd_pool_stats", "ikd_get_counters", "ikd_device_bytes", "ikd_acquire_removed_points", "ikd_set_tuning", "ikd_dump_nodes", "ikd_last_error",
    "ikd_batch_create", "ikd_batch_destroy", "ikd_batch_nearest_search_device", "ikd_batch_synchronize",
    "ikd_batch_stream", "ikd_batch_set_tuning", "ikd_batch_last_error", "ikd_batch_radius_search_device",
    "ikd_batch_nearest_search_noalloc", "ikd_batch_knn_graph_create", "ikd_batch_knn_graph_launch",
    "ikd_batch_knn_graph_destroy",
    "ikd_now_ns", "ikd_server_create", "ikd_server_add_map", "ikd_server_submit", "ikd_server_wait",
    "ikd_server_pending", "ikd_server_destroy", "ikd_server_set_maintenance", "ikd_server_stats",
    "ikd_server_memory", "ikd_host_alloc", "ikd_host_free",
    "ikd_rebuild_deferred", "ikd_list_payload_ids", "ikd_relabel_payload_ids",
)
IKD_TUNE_QUERY_ORDER = 1
IKD_TUNE_KNN_STACK = 2
IKD_TUNE_LANES_PER_QUERY = 3
IKD_TUNE_SMALL_BUILD = 4
IKD_TUNE_DEFER_REBUILD = 5
IKD_TUNE_FOLD_LANES = 6
IKD_TUNE_RECORD_REMOVED = 7
IKD_TUNE_KNN_STAGE = 8
IKD_TUNE_KNN_COMPACT = 9
IKD_NOALLOC_MAX_TREES = 128
IKD_NOALLOC_MAX_QUERIES = 1 << 20


class Counters(ctypes.Structure):
    """ikd_counters_t: a tree's event counts since creation."""
    _fields_ = [(name, ctypes.c_int64) for name in
                ("rebuilds", "rebuilt_points", "compactions", "pool_grows", "scratch_grows", "host_waits")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class IkdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ikd error {code}: {msg}")
        self.code = code


def build_library(jobs: int = 8) -> str:
    """Compile libikd_hip.so for gfx950 with hipcc (csrc/Makefile)."""
    subprocess.run(["make", "-C", CSRC, f"-j{jobs}", "-B"], check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


_libs = {}


def load_library(path: str | None = None):
    """Load the HIP library (or, for debugging, another build of the same
    sources at `path`); raises ImportError if it has not been built."""
    path = path or LIB_PATH
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with build_library() or `make -C {CSRC}`")
    lib = ctypes.CDLL(path)
    vp, i64, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_double
    sig = {
        "ikd_create": [f32, f32, f32, ctypes.POINTER(vp)],
        "ikd_destroy": [vp],
        "ikd_set_delete_criterion_param": [vp, f32],
        "ikd_set_balance_criterion_param": [vp, f32],
        "ikd_set_downsample_param": [vp, f32],
        "ikd_initialize": [vp, f32, f32, f32],
        "ikd_build": [vp, vp, vp, i64],
        "ikd_nearest_search": [vp, vp, i64, i32, f64, vp, vp, vp, vp],
        "ikd_nearest_search_device": [vp, vp, i64, i32, f64, vp, vp, vp, vp],
        "ikd_nearest_search_counters": [vp, vp, i64, i32, f64, vp, vp],
        "ikd_box_search": [vp, vp, i64, vp, vp, vp, i64],
        "ikd_radius_search": [vp, vp, vp, i64, vp, vp, vp, i64],
        "ikd_add_points": [vp, vp, vp, i64, i32, vp],
        "ikd_add_points_device": [vp, vp, vp, i64, i32, vp],
        "ikd_delete_points": [vp, vp, i64],
        "ikd_delete_point_boxes": [vp, vp, i64, vp],
        "ikd_add_point_boxes": [vp, vp, i64],
        "ikd_size": [vp, vp],
        "ikd_validnum": [vp, vp],
        "ikd_tree_range": [vp, vp],
        "ikd_root_alpha": [vp, vp, vp],
        "ikd_flatten": [vp, vp, vp, i64, vp],
        "ikd_synchronize": [vp],
        "ikd_stream": [vp, ctypes.POINTER(vp)],
        "ikd_pool_stats": [vp, vp, vp, vp, vp],
        "ikd_get_counters": [vp, vp],
        "ikd_device_bytes": [vp, ctypes.POINTER(ctypes.c_int64)],
        "ikd_acquire_removed_points": [vp, vp, vp, i64, vp],
        "ikd_set_tuning": [vp, i32, i32],
        "ikd_dump_nodes": [vp, vp, i64, vp],
        "ikd_rebuild_deferred": [vp, i64, vp],
        "ikd_list_payload_ids": [vp, vp, i64, vp],
        "ikd_relabel_payload_ids": [vp, vp, i64],
        "ikd_batch_create": [ctypes.POINTER(vp)],
        "ikd_batch_destroy": [vp],
        "ikd_batch_nearest_search_device": [vp, vp, i32, vp, vp, i64, vp, i32, vp, f64, vp, vp, vp, vp],
        "ikd_batch_nearest_search_noalloc": [vp, vp, i32, vp, vp, i64, vp, i32, vp, f64, vp, vp, vp, vp],
        "ikd_batch_knn_graph_create": [vp, vp, i32, vp, vp, i64, vp, i32, vp, f64, vp, vp, vp, vp, ctypes.POINTER(vp)],
        "ikd_batch_knn_graph_launch": [vp],
        "ikd_batch_knn_graph_destroy": [vp],
        "ikd_batch_synchronize": [vp],
        "ikd_batch_radius_search_device": [vp, vp, i32, vp, vp, vp, i64, vp, vp, vp, i64, vp],
        "ikd_batch_stream": [vp, ctypes.POINTER(vp)],
        "ikd_batch_set_tuning": [vp, i32, i32],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = i32
    lib.ikd_last_error.argtypes = [vp]
    lib.ikd_last_error.restype = ctypes.c_char_p
    lib.ikd_batch_last_error.argtypes = [vp]
    lib.ikd_batch_last_error.restype = ctypes.c_char_p
    _libs[path] = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ctypes.c_void_p)
    return ctypes.c_void_p(int(a))  # a raw device address


def _pts(points) -> np.ndarray:
    return np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)


def _tensor_ptr(t):
    """Device address of a torch tensor (contiguous, on the GPU)."""
    if not t.is_contiguous():
        raise ValueError("device tensors must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class KD_TREE:
    """``KD_TREE<PointType>`` (ikd-Tree/ikd_Tree.h:59-249) on one MI355X."""

    def __init__(self, delete_param: float = 0.5, balance_param: float = 0.6, box_length: float = 0.2, lib=None):
        self._lib = lib if lib is not None else load_library()
        h = ctypes.c_void_p()
        self._check(self._lib.ikd_create(delete_param, balance_param, box_length, ctypes.byref(h)), handle=None)
        self._h = h

    # ------------------------------------------------------------ plumbing
    def _check(self, rc: int, handle="self"):
        if rc != IKD_OK:
            h = self._h if handle == "self" else None
            msg = self._lib.ikd_last_error(h).decode() if h else "create failed"
            raise IkdError(rc, msg)
        return rc

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ikd_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ parameters
    def Set_delete_criterion_param(self, delete_param: float):
        self._check(self._lib.ikd_set_delete_criterion_param(self._h, delete_param))

    def Set_balance_criterion_param(self, balance_param: float):
        self._check(self._lib.ikd_set_balance_criterion_param(self._h, balance_param))

    def set_downsample_param(self, box_length: float):
        self._check(self._lib.ikd_set_downsample_param(self._h, box_length))

    def InitializeKDTree(self, delete_param: float = 0.5, balance_param: float = 0.7, box_length: float = 0.2):
        self._check(self._lib.ikd_initialize(self._h, delete_param, balance_param, box_length))

    # ------------------------------------------------------------ build
    def Build(self, points, ids=None):
        p = _pts(points)
        i = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        self._check(self._lib.ikd_build(self._h, _ptr(p), _ptr(i), p.shape[0]))

    # ------------------------------------------------------------ kNN
    def knn(self, queries, k: int, max_dist: float = float("inf")):
        """Batched Nearest_Search: (xyz[nq,k,3], d2[nq,k], ids[nq,k], counts[nq])."""
        q = _pts(queries)
        nq = q.shape[0]
        xyz = np.zeros((nq, k, 3), np.float32)
        d2 = np.zeros((nq, k), np.float32)
        ids = np.zeros((nq, k), np.uint32)
        cnt = np.zeros(nq, np.int32)
        self._check(self._lib.ikd_nearest_search(self._h, _ptr(q), nq, k, float(max_dist), _ptr(xyz), _ptr(d2),
                                                 _ptr(ids), _ptr(cnt)))
        return xyz, d2, ids, cnt

    def Nearest_Search(self, point, k_nearest: int, max_dist: float = float("inf")):
        """One query (cpp:366-397): (Nearest_Points[n,3], Point_Distance[n]) ascending."""
        xyz, d2, _, cnt = self.knn(np.asarray(point, np.float32).reshape(1, 3), k_nearest, max_dist)
        n = int(cnt[0])
        return xyz[0, :n].copy(), d2[0, :n].copy()

    def knn_counters(self, queries, k: int, max_dist: float = float("inf")):
        """Per-query nodes visited and child boxes evaluated (instrumentation)."""
        q = _pts(queries)
        nq = q.shape[0]
        v = np.zeros(nq, np.uint32)
        b = np.zeros(nq, np.uint32)
        self._check(self._lib.ikd_nearest_search_counters(self._h, _ptr(q), nq, k, float(max_dist), _ptr(v), _ptr(b)))
        return v, b

    def knn_device(self, q, k: int, max_dist: float, out_xyz, out_d2, out_ids, out_counts):
        """Stream-ordered kNN on device tensors (torch, contiguous): q[nq,3] float32,
        out_xyz[nq,k,3], out_d2[nq,k] float32, out_ids[nq,k] int32/uint32, out_counts[nq] int32."""
        self._check(self._lib.ikd_nearest_search_device(self._h, _tensor_ptr(q), q.shape[0], k, float(max_dist),
                                                        _tensor_ptr(out_xyz), _tensor_ptr(out_d2),
                                                        _tensor_ptr(out_ids), _tensor_ptr(out_counts)))

    # ------------------------------------------------------------ range queries
    def _range(self, fn, args, n):
        off = np.zeros(n + 1, np.int64)
        self._check(fn(self._h, *args, n, _ptr(off), None, None, 0))
        total = int(off[n])
        xyz = np.zeros((max(total, 1), 3), np.float32)
        ids = np.zeros(max(total, 1), np.uint32)
        self._check(fn(self._h, *args, n, _ptr(off), _ptr(xyz), _ptr(ids), total))
        return off, xyz[:total], ids[:total]

    def box_search_batch(self, boxes):
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        return self._range(self._lib.ikd_box_search, (_ptr(b),), b.shape[0])

    def Box_Search(self, box):
        """Live points in the half-open box (cpp:399-404)."""
        _, xyz, _ = self.box_search_batch(np.asarray(box, np.float32).reshape(1, 6))
        return xyz

    def radius_search_batch(self, centers, radii):
        c = _pts(centers)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(radii, np.float32), (c.shape[0],)))
        return self._range(self._lib.ikd_radius_search, (_ptr(c), _ptr(r)), c.shape[0])

    def Radius_Search(self, point, radius: float):
        """Live points within radius (cpp:406-411)."""
        _, xyz, _ = self.radius_search_batch(np.asarray(point, np.float32).reshape(1, 3), [radius])
        return xyz

    # ------------------------------------------------------------ updates
    def Add_Points(self, points, downsample_on: bool, ids=None) -> int:
        p = _pts(points)
        i = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        added = ctypes.c_int32(0)
        self._check(self._lib.ikd_add_points(self._h, _ptr(p), _ptr(i), p.shape[0], 1 if downsample_on else 0,
                                             ctypes.byref(added)))
        return int(added.value)

    def add_points_device(self, xyz, downsample_on: bool, ids=None) -> int:
        """Add_Points with a device tensor xyz[n,3] (and optional ids[n])."""
        added = ctypes.c_int32(0)
        self._check(self._lib.ikd_add_points_device(self._h, _tensor_ptr(xyz),
                                                    None if ids is None else _tensor_ptr(ids), xyz.shape[0],
                                                    1 if downsample_on else 0, ctypes.byref(added)))
        return int(added.value)

    def Delete_Points(self, points):
        p = _pts(points)
        self._check(self._lib.ikd_delete_points(self._h, _ptr(p), p.shape[0]))

    def Delete_Point_Boxes(self, boxes) -> int:
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        out = ctypes.c_int32(0)
        self._check(self._lib.ikd_delete_point_boxes(self._h, _ptr(b), b.shape[0], ctypes.byref(out)))
        return int(out.value)

    def Add_Point_Boxes(self, boxes):
        """Add_Point_Boxes (ikd_Tree.cpp:491-511): restore the deleted points the
        boxes reach that the downsample did not delete (no return value)."""
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self._check(self._lib.ikd_add_point_boxes(self._h, _ptr(b), b.shape[0]))

    # ------------------------------------------------------------ accessors
    def size(self) -> int:
        v = ctypes.c_int64(0)
        self._check(self._lib.ikd_size(self._h, ctypes.byref(v)))
        return int(v.value)

    def validnum(self) -> int:
        v = ctypes.c_int64(0)
        self._check(self._lib.ikd_validnum(self._h, ctypes.byref(v)))
        return int(v.value)

    def tree_range(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        self._check(self._lib.ikd_tree_range(self._h, _ptr(out)))
        return out

    def root_alpha(self):
        a = ctypes.c_float(0)
        d = ctypes.c_float(0)
        self._check(self._lib.ikd_root_alpha(self._h, ctypes.byref(a), ctypes.byref(d)))
        return float(a.value), float(d.value)

    def flatten(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.ikd_flatten(self._h, None, None, 0, ctypes.byref(n)))
        cnt = int(n.value)
        xyz = np.zeros((max(cnt, 1), 3), np.float32)
        ids = np.zeros(max(cnt, 1), np.uint32)
        self._check(self._lib.ikd_flatten(self._h, _ptr(xyz), _ptr(ids), cnt, ctypes.byref(n)))
        return xyz[:cnt], ids[:cnt]

    def pool_stats(self):
        used = ctypes.c_int64(0)
        cap = ctypes.c_int64(0)
        depth = ctypes.c_int32(0)
        reb = ctypes.c_int64(0)
        self._check(self._lib.ikd_pool_stats(self._h, ctypes.byref(used), ctypes.byref(cap), ctypes.byref(depth),
                                             ctypes.byref(reb)))
        return {"used_nodes": used.value, "capacity_nodes": cap.value, "max_depth": depth.value,
                "rebuilds": reb.value}

    def acquire_removed_points(self):
        """acquire_removed_points (ikd_Tree.cpp:558-571): the points rebuilds
        removed since the last call (recorded while IKD_TUNE_RECORD_REMOVED is
        1), as (xyz, ids); clears the record."""
        n = ctypes.c_int64(0)
        rc = self._lib.ikd_acquire_removed_points(self._h, None, None, 0, ctypes.byref(n))
        if rc not in (IKD_OK, IKD_ERR_CAPACITY):
            self._check(rc)
        cnt = n.value
        xyz = np.zeros((max(cnt, 1), 3), np.float32)
        ids = np.zeros(max(cnt, 1), np.uint32)
        if cnt > 0:
            self._check(self._lib.ikd_acquire_removed_points(self._h, _ptr(xyz), _ptr(ids), cnt, ctypes.byref(n)))
        return xyz[:cnt], ids[:cnt]

    def counters(self):
        c = Counters()
        self._check(self._lib.ikd_get_counters(self._h, ctypes.byref(c)))
        return c.as_dict()

    def device_bytes(self) -> int:
        """Device memory the tree holds (node pool, scratch, records)."""
        v = ctypes.c_int64(0)
        self._check(self._lib.ikd_device_bytes(self._h, ctypes.byref(v)))
        return int(v.value)

    def set_tuning(self, key: int, value: int):
        self._check(self._lib.ikd_set_tuning(self._h, key, value))

    def rebuild_deferred(self, max_points: int = 0) -> int:
        """Run the deferred rebuilds still due (see IKD_TUNE_DEFER_REBUILD); returns how many ran."""
        n = ctypes.c_int32(0)
        self._check(self._lib.ikd_rebuild_deferred(self._h, int(max_points), ctypes.byref(n)))
        return int(n.value)

    def list_payload_ids(self):
        """Ids the tree can still return (every node and the removed-point record), ascending."""
        n = ctypes.c_int64(0)
        self._check(self._lib.ikd_list_payload_ids(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint32)
        self._check(self._lib.ikd_list_payload_ids(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def relabel_payload_ids(self, old_ids):
        """Id old_ids[i] becomes i (old_ids strictly ascending)."""
        o = np.ascontiguousarray(old_ids, np.uint32)
        self._check(self._lib.ikd_relabel_payload_ids(self._h, _ptr(o), o.shape[0]))

    def dump_nodes(self):
        """Pre-order rows (oracle DUMP_DTYPE layout) of the device tree."""
        n = ctypes.c_int64(0)
        self._check(self._lib.ikd_dump_nodes(self._h, None, 0, ctypes.byref(n)))
        dt = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("id", "<u4"), ("axis", "<i4"),
                       ("tree_size", "<i4"), ("invalid", "<i4"), ("flags", "<i4"), ("range", "<f4", (6,)),
                       ("left", "<i4"), ("right", "<i4")])
        rows = np.zeros(max(n.value, 1), dt)
        self._check(self._lib.ikd_dump_nodes(self._h, _ptr(rows), n.value, ctypes.byref(n)))
        return rows[:n.value]

    def synchronize(self):
        self._check(self._lib.ikd_synchronize(self._h))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(self._lib.ikd_stream(self._h, ctypes.byref(s)))
        return s.value or 0


class KnnGraph:
    """A multi-tree kNN launch recorded into a HIP graph (Batch.knn_graph)."""

    def __init__(self, batch, h, trees):
        self._batch = batch
        self._h = h
        self._trees = trees  # the trees must outlive the recording

    def launch(self):
        self._batch._check(self._batch._lib.ikd_batch_knn_graph_launch(self._h))

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._batch._lib.ikd_batch_knn_graph_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many trees' kNN queries in one launch (the batched Nearest_Search).

    ``nearest_search_device(trees, qtree, q, k_max, k=None, max_dist=inf, outs)``
    runs query i against ``trees[qtree[i]]`` with its own k (``k`` an int32
    tensor, or None for k_max) and max_dist (a float64 tensor, or one float).
    Results are identical to one ``KD_TREE.knn_device`` call per tree."""

    def __init__(self, lib=None):
        self._lib = lib if lib is not None else load_library()
        h = ctypes.c_void_p()
        rc = self._lib.ikd_batch_create(ctypes.byref(h))
        if rc != IKD_OK:
            raise IkdError(rc, "batch create failed")
        self._h = h

    def _check(self, rc):
        if rc != IKD_OK:
            raise IkdError(rc, self._lib.ikd_batch_last_error(self._h).decode())

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ikd_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, key: int, value: int):
        self._check(self._lib.ikd_batch_set_tuning(self._h, key, value))

    def nearest_search_device(self, trees, qtree, q, k_max: int, out_xyz, out_d2, out_ids, out_counts, k=None,
                              max_dist=float("inf")):
        handles = (ctypes.c_void_p * len(trees))(*[t._h.value for t in trees])
        per_q = not isinstance(max_dist, (int, float))
        self._check(self._lib.ikd_batch_nearest_search_device(
            self._h, handles, len(trees), _tensor_ptr(qtree), _tensor_ptr(q), q.shape[0],
            None if k is None else _tensor_ptr(k), int(k_max), _tensor_ptr(max_dist) if per_q else None,
            float("inf") if per_q else float(max_dist), _tensor_ptr(out_xyz), _tensor_ptr(out_d2),
            _tensor_ptr(out_ids), _tensor_ptr(out_counts)))

    def nearest_search_noalloc(self, trees, qtree, q, k_max: int, out_xyz, out_d2, out_ids, out_counts, k=None,
                               max_dist=float("inf")):
        """nearest_search_device without any device allocation (the library's
        static scratch array; at most IKD_NOALLOC_MAX_TREES trees, and with
        per-query max_dist at most IKD_NOALLOC_MAX_QUERIES queries)."""
        handles = (ctypes.c_void_p * len(trees))(*[t._h.value for t in trees])
        per_q = not isinstance(max_dist, (int, float))
        self._check(self._lib.ikd_batch_nearest_search_noalloc(
            self._h, handles, len(trees), _tensor_ptr(qtree), _tensor_ptr(q), q.shape[0],
            None if k is None else _tensor_ptr(k), int(k_max), _tensor_ptr(max_dist) if per_q else None,
            float("inf") if per_q else float(max_dist), _tensor_ptr(out_xyz), _tensor_ptr(out_d2),
            _tensor_ptr(out_ids), _tensor_ptr(out_counts)))

    def knn_graph(self, trees, qtree, q, k_max: int, out_xyz, out_d2, out_ids, out_counts, k=None,
                  max_dist=float("inf")):
        """Records nearest_search_device with these buffers into a HIP graph;
        the returned KnnGraph's launch() replays it (refill the buffers between
        replays as needed)."""
        handles = (ctypes.c_void_p * len(trees))(*[t._h.value for t in trees])
        per_q = not isinstance(max_dist, (int, float))
        g = ctypes.c_void_p()
        self._check(self._lib.ikd_batch_knn_graph_create(
            self._h, handles, len(trees), _tensor_ptr(qtree), _tensor_ptr(q), q.shape[0],
            None if k is None else _tensor_ptr(k), int(k_max), _tensor_ptr(max_dist) if per_q else None,
            float("inf") if per_q else float(max_dist), _tensor_ptr(out_xyz), _tensor_ptr(out_d2),
            _tensor_ptr(out_ids), _tensor_ptr(out_counts), ctypes.byref(g)))
        return KnnGraph(self, g, list(trees))

    def radius_search_device(self, trees, qtree, centers, radii, out_offsets, out_xyz=None, out_ids=None):
        """Radius_Search over many trees; returns the total result count.  With
        out_xyz/out_ids (device tensors of at least total rows) it also fills them."""
        handles = (ctypes.c_void_p * len(trees))(*[t._h.value for t in trees])
        total = ctypes.c_int64(0)
        cap = 0 if out_xyz is None else int(out_xyz.shape[0])
        self._check(self._lib.ikd_batch_radius_search_device(
            self._h, handles, len(trees), _tensor_ptr(qtree), _tensor_ptr(centers), _tensor_ptr(radii),
            centers.shape[0], _tensor_ptr(out_offsets), None if out_xyz is None else _tensor_ptr(out_xyz),
            None if out_ids is None else _tensor_ptr(out_ids), cap, ctypes.byref(total)))
        return int(total.value)

    def synchronize(self):
        self._check(self._lib.ikd_batch_synchronize(self._h))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(self._lib.ikd_batch_stream(self._h, ctypes.byref(s)))
        return s.value or 0
