"""Writes c1_uniform_100k.npz: the golden vectors of the small plain-method
workload (Build of 100k points uniform in [-50, 50)^3, k=5 Nearest_Search of
10k uniform queries, Delete_Point_Boxes of 10 boxes with sides 5-20 m, the
same queries again), computed by numpy brute force over the live point set —
independent of the oracle and of the HIP library.

Inputs come from the seeded generator in ikd-tree_amd/csrc/scene.cpp (seed
1001: points stream 0, boxes stream 1, queries stream 2); the fixture stores
their SHA-256 so a test can tell a generator change from a result change.
Expected results are kept for the first 2000 queries (d2 and ids in
(d2, x, y, z, id) order) plus the delete count and the live count.

Run from the repository root:  python tests/golden/make_c1_fixture.py
"""
import hashlib
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from common import brute_knn, in_box  # noqa: E402

N_POINTS, N_QUERIES, N_BOXES, K, KEEP = 100_000, 10_000, 10, 5, 2000


def inputs():
    scenes = importlib.import_module("ikd-tree_amd.scenes")
    pts = scenes.uniform(1001, N_POINTS, -50.0, 50.0, stream=0)
    boxes = scenes.boxes(1001, N_BOXES, -50.0, 50.0, 5.0, 20.0, stream=1)
    q = scenes.uniform(1001, N_QUERIES, -50.0, 50.0, stream=2)
    return pts, boxes, q


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    pts, boxes, q = inputs()
    ids = np.arange(N_POINTS, dtype=np.uint32)
    d0, i0, c0 = brute_knn(pts, ids, q[:KEEP], K, tie="xyz")
    dead = np.zeros(N_POINTS, bool)
    for b in boxes:
        dead |= in_box(pts, b)
    live = ~dead
    d1, i1, c1 = brute_knn(pts[live], ids[live], q[:KEEP], K, tie="xyz")
    np.savez_compressed(os.path.join(HERE, "c1_uniform_100k.npz"), input_sha256=np.array(sha(pts, boxes, q)),
                        d2_before=d0, ids_before=i0, d2_after=d1, ids_after=i1,
                        deleted=np.int64(dead.sum()), validnum_after=np.int64(live.sum()))
    print("deleted", int(dead.sum()), "live", int(live.sum()))


if __name__ == "__main__":
    main()
