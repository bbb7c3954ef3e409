"""The small plain-method workload against its committed golden vectors
(tests/golden/c1_uniform_100k.npz, numpy brute force; see make_c1_fixture.py):
Build of 100k uniform points, k=5 Nearest_Search, Delete_Point_Boxes of 10
boxes, the same queries again.  The CPU test pins the oracle; the GPU test
checks the HIP path through the C ABI, bit for bit."""
import importlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_c1_fixture import K, KEEP, inputs, sha  # noqa: E402

FIX = os.path.join(HERE, "golden", "c1_uniform_100k.npz")


@pytest.fixture(scope="module")
def golden():
    g = np.load(FIX)
    pts, boxes, q = inputs()
    assert str(g["input_sha256"]) == sha(pts, boxes, q), "input generator changed"
    return g, pts, boxes, q


def _check(d, i, c, gd, gi):
    assert np.all(c == K)
    assert np.array_equal(d.view(np.uint32), gd.view(np.uint32))
    assert np.array_equal(i, gi)


def test_oracle_matches_golden(golden):
    from oracle import OracleTree
    g, pts, boxes, q = golden
    o = OracleTree(0.5, 0.6, 0.2)
    o.Build(pts)
    _, d, i, c = o.knn(q[:KEEP], K)
    _check(d, i, c, g["d2_before"], g["ids_before"])
    assert o.Delete_Point_Boxes(boxes) == int(g["deleted"])
    assert o.validnum() == int(g["validnum_after"])
    _, d, i, c = o.knn(q[:KEEP], K)
    _check(d, i, c, g["d2_after"], g["ids_after"])


@pytest.mark.gpu
def test_hip_matches_golden(golden):
    ikd = importlib.import_module("ikd-tree_amd")
    from oracle import OracleTree
    g, pts, boxes, q = golden
    t = ikd.KD_TREE(0.5, 0.6, 0.2)
    t.Build(pts)
    _, d, i, c = t.knn(q, K)
    _check(d[:KEEP], i[:KEEP], c[:KEEP], g["d2_before"], g["ids_before"])
    o = OracleTree(0.5, 0.6, 0.2)
    o.Build(pts)
    _, od, oi, oc = o.knn(q, K)  # all 10k queries against the oracle
    assert np.array_equal(c, oc) and np.array_equal(d.view(np.uint32), od.view(np.uint32))
    assert t.Delete_Point_Boxes(boxes) == int(g["deleted"])
    assert t.validnum() == int(g["validnum_after"])
    _, d, i, c = t.knn(q, K)
    _check(d[:KEEP], i[:KEEP], c[:KEEP], g["d2_after"], g["ids_after"])
