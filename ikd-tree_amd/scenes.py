"""Deterministic synthetic workloads (wraps csrc/scene.cpp, built as libikd_scene.so).

- ``uniform(seed, n, lo, hi)``: uniform points in a cube (the 100k-point
  Build + 10k kNN + 10 box deletes workload uses seed 1001).
- ``boxes(seed, n, lo, hi, smin, smax)``: random half-open boxes.
- ``MappingScene``: the LiDAR-like incremental mapping workload: ground plane
  plus 300 boxes in a 400 x 400 m area, a 32-beam x 1800-azimuth scan per
  frame from a sensor moving 1 m per frame on a circle of radius 80 m
  (~30-40k points per frame), seed 1002.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libikd_scene.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError(f"{_LIB} is missing: run `make -C {os.path.join(_HERE, 'csrc')}`")
        lib = ctypes.CDLL(_LIB)
        u64, i64, d, vp = ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        lib.scene_uniform.argtypes = [u64, u64, i64, d, d, vp]
        lib.scene_boxes.argtypes = [u64, u64, i64, d, d, d, d, vp]
        lib.scene_obstacles.argtypes = [u64, i64, d, vp]
        lib.scene_pose.argtypes = [i64, d, vp]
        lib.scene_scan.argtypes = [vp, i64, i64, u64, d, vp]
        lib.scene_scan.restype = i64
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def uniform(seed: int, n: int, lo: float = -50.0, hi: float = 50.0, stream: int = 0) -> np.ndarray:
    out = np.empty((n, 3), np.float32)
    _load().scene_uniform(seed, stream, n, lo, hi, _p(out))
    return out


def boxes(seed: int, n: int, lo: float = -50.0, hi: float = 50.0, smin: float = 5.0, smax: float = 20.0,
          stream: int = 1) -> np.ndarray:
    out = np.empty((n, 6), np.float32)
    _load().scene_boxes(seed, stream, n, lo, hi, smin, smax, _p(out))
    return out


class MappingScene:
    """Per-frame scans of the incremental mapping workload."""

    N_OBSTACLES = 300
    PATH_RADIUS = 80.0
    MAX_POINTS = 32 * 1800

    def __init__(self, seed: int = 1002, n_obstacles: int = N_OBSTACLES):
        self.seed = seed
        self.obstacles = np.empty((n_obstacles, 6), np.float32)
        _load().scene_obstacles(seed, n_obstacles, self.PATH_RADIUS, _p(self.obstacles))

    def pose(self, frame: int) -> np.ndarray:
        out = np.empty(3, np.float64)
        _load().scene_pose(frame, self.PATH_RADIUS, _p(out))
        return out

    def scan(self, frame: int) -> np.ndarray:
        buf = np.empty((self.MAX_POINTS, 3), np.float32)
        n = _load().scene_scan(_p(self.obstacles), self.obstacles.shape[0], frame, self.seed, self.PATH_RADIUS,
                               _p(buf))
        return buf[:n].copy()
