"""ctypes wrapper around the CPU oracle (oracle/ikd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg as the checker / CPU baseline.  The
product path (the ``ikd-tree_amd`` package and its HIP library) never imports
it.

The oracle restates the reference ``KD_TREE<PointType>`` of
hku-mars/ikd-Tree (ikd-Tree/ikd_Tree.cpp) on the CPU: pointer tree, lazy
labels, nth_element build, bounded heap, synchronous rebuilds.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libikd_oracle.so")
_lib = None

DUMP_DTYPE = np.dtype(
    [
        ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("id", "<u4"),
        ("axis", "<i4"), ("tree_size", "<i4"), ("invalid", "<i4"), ("flags", "<i4"),
        ("range", "<f4", (6,)), ("left", "<i4"), ("right", "<i4"),
    ]
)


def build_library(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "libikd_oracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build_library()
    lib = ctypes.CDLL(_LIB_PATH)
    vp, f32, i64, i32 = ctypes.c_void_p, ctypes.c_float, ctypes.c_int64, ctypes.c_int
    lib.ora_create.restype = vp
    lib.ora_create.argtypes = [f32, f32, f32]
    lib.ora_destroy.argtypes = [vp]
    lib.ora_set_params.argtypes = [vp, f32, f32, f32]
    lib.ora_build.argtypes = [vp, vp, vp, i64]
    lib.ora_knn.restype = i32
    lib.ora_knn.argtypes = [vp, vp, i64, i32, ctypes.c_double, vp, vp, vp, vp, vp, vp, i32]
    lib.ora_box_search.restype = i64
    lib.ora_box_search.argtypes = [vp, vp, vp, vp, i64]
    lib.ora_radius_search.restype = i64
    lib.ora_radius_search.argtypes = [vp, vp, f32, vp, vp, i64]
    lib.ora_add_points.restype = i32
    lib.ora_add_points.argtypes = [vp, vp, vp, i64, i32]
    lib.ora_delete_points.argtypes = [vp, vp, i64]
    lib.ora_delete_point_boxes.restype = i32
    lib.ora_delete_point_boxes.argtypes = [vp, vp, i64]
    lib.ora_add_point_boxes.argtypes = [vp, vp, i64]
    lib.ora_size.restype = i64
    lib.ora_size.argtypes = [vp]
    lib.ora_validnum.restype = i64
    lib.ora_validnum.argtypes = [vp]
    lib.ora_tree_range.argtypes = [vp, vp]
    lib.ora_root_alpha.argtypes = [vp, vp, vp]
    lib.ora_flatten.restype = i64
    lib.ora_flatten.argtypes = [vp, vp, vp, i64]
    lib.ora_dump.restype = i64
    lib.ora_dump.argtypes = [vp, vp, i64]
    lib.ora_acquire_removed.restype = i64
    lib.ora_acquire_removed.argtypes = [vp, vp, vp, i64]
    lib.ora_dump_row_bytes.restype = i32
    assert lib.ora_dump_row_bytes() == DUMP_DTYPE.itemsize
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _pts(points) -> np.ndarray:
    a = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    return a


class OracleTree:
    """CPU restatement of ``KD_TREE<PointType>`` (ikd-Tree/ikd_Tree.h:225-248)."""

    def __init__(self, delete_param: float = 0.5, balance_param: float = 0.6, box_length: float = 0.2):
        self._lib = _load()
        self._h = self._lib.ora_create(delete_param, balance_param, box_length)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ora_destroy(h)
            self._h = None

    def set_params(self, delete_param: float, balance_param: float, box_length: float) -> None:
        self._lib.ora_set_params(self._h, delete_param, balance_param, box_length)

    def Build(self, points, ids=None) -> None:
        p = _pts(points)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        self._lib.ora_build(self._h, _ptr(p), _ptr(i), p.shape[0])

    def knn(self, queries, k: int, max_dist: float = float("inf"), counters: bool = False, threads: int = 1):
        """Batched Nearest_Search: returns (xyz[nq,k,3], d2[nq,k], ids[nq,k], counts[nq])."""
        q = _pts(queries)
        nq = q.shape[0]
        xyz = np.zeros((nq, k, 3), np.float32)
        d2 = np.zeros((nq, k), np.float32)
        ids = np.zeros((nq, k), np.uint32)
        cnt = np.zeros(nq, np.int32)
        vis = np.zeros(nq, np.int64) if counters else None
        cb = np.zeros(nq, np.int64) if counters else None
        rc = self._lib.ora_knn(self._h, _ptr(q), nq, k, float(max_dist), _ptr(xyz), _ptr(d2), _ptr(ids),
                               _ptr(cnt), _ptr(vis), _ptr(cb), threads)
        if rc != 0:
            raise ValueError("oracle knn rejected k=%d" % k)
        if counters:
            return xyz, d2, ids, cnt, vis, cb
        return xyz, d2, ids, cnt

    def Nearest_Search(self, point, k: int, max_dist: float = float("inf")):
        xyz, d2, ids, cnt = self.knn(np.asarray(point, np.float32).reshape(1, 3), k, max_dist)
        n = int(cnt[0])
        return xyz[0, :n], d2[0, :n], ids[0, :n]

    def _collect(self, fn, *args):
        n = fn(self._h, *args, None, None, 0)
        xyz = np.zeros((max(n, 1), 3), np.float32)
        ids = np.zeros(max(n, 1), np.uint32)
        n2 = fn(self._h, *args, _ptr(xyz), _ptr(ids), n)
        assert n2 == n
        return xyz[:n], ids[:n]

    def Box_Search(self, box):
        b = np.ascontiguousarray(box, np.float32).reshape(6)
        return self._collect(self._lib.ora_box_search, _ptr(b))

    def Radius_Search(self, point, radius: float):
        q = np.ascontiguousarray(point, np.float32).reshape(3)
        return self._collect(self._lib.ora_radius_search, _ptr(q), ctypes.c_float(radius))

    def Add_Points(self, points, downsample_on: bool, ids=None) -> int:
        p = _pts(points)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        return int(self._lib.ora_add_points(self._h, _ptr(p), _ptr(i), p.shape[0], 1 if downsample_on else 0))

    def Delete_Points(self, points) -> None:
        p = _pts(points)
        self._lib.ora_delete_points(self._h, _ptr(p), p.shape[0])

    def Delete_Point_Boxes(self, boxes) -> int:
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        return int(self._lib.ora_delete_point_boxes(self._h, _ptr(b), b.shape[0]))

    def Add_Point_Boxes(self, boxes) -> None:
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self._lib.ora_add_point_boxes(self._h, _ptr(b), b.shape[0])

    def size(self) -> int:
        return int(self._lib.ora_size(self._h))

    def validnum(self) -> int:
        return int(self._lib.ora_validnum(self._h))

    def tree_range(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        self._lib.ora_tree_range(self._h, _ptr(out))
        return out

    def root_alpha(self):
        a = np.zeros(1, np.float32)
        d = np.zeros(1, np.float32)
        self._lib.ora_root_alpha(self._h, _ptr(a), _ptr(d))
        return float(a[0]), float(d[0])

    def flatten(self):
        return self._collect(self._lib.ora_flatten)

    def acquire_removed_points(self):
        """acquire_removed_points (cpp:558-571): points rebuilds removed since
        the last call, deleted but not by the downsample; clears the record."""
        n = int(self._lib.ora_acquire_removed(self._h, None, None, 0))
        xyz = np.zeros((max(n, 1), 3), np.float32)
        ids = np.zeros(max(n, 1), np.uint32)
        got = int(self._lib.ora_acquire_removed(self._h, _ptr(xyz), _ptr(ids), n))
        assert got == n
        return xyz[:n], ids[:n]

    def dump(self) -> np.ndarray:
        n = self._lib.ora_dump(self._h, None, 0)
        rows = np.zeros(max(n, 1), DUMP_DTYPE)
        self._lib.ora_dump(self._h, _ptr(rows), n)
        return rows[:n]
