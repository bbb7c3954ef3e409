"""GPU: box/radius search and sliding-window trims on a large LiDAR-like map.

A map of raw scans of the mapping scene (1M points here), then box queries
with log-uniform sides 1-20 m and radius queries with r 0.2-3 m centred on
live points, and sliding-window trims (six half-open slabs around the moving
sensor) that drive alpha_del / alpha_bal subtree rebuilds on the device.

Bars, size-independent: box results equal the half-open filter over the live
set exactly; radius results equal brute force outside the sqrt band; each
trim's return equals the brute-force count of live points it covers; the live
set after the trims is the set algebra of the deletions; kNN after the trims
equals brute force.
"""
import importlib

import numpy as np
import pytest

from common import assert_same_multiset, brute_knn, brute_radius, in_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ikd():
    return importlib.import_module("ikd-tree_amd")


@pytest.fixture(scope="module")
def scenes():
    return importlib.import_module("ikd-tree_amd.scenes")


def _slabs(p, half):
    big = 1e7
    lo, hi = p - half, p + half
    return np.array([
        [-big, -big, -big, lo[0], big, big], [hi[0], -big, -big, big, big, big],
        [-big, -big, -big, big, lo[1], big], [-big, hi[1], -big, big, big, big],
        [-big, -big, -big, big, big, lo[2]], [-big, -big, hi[2], big, big, big],
    ], np.float32)


def test_large_map_queries_and_trims(ikd, scenes):
    sc = scenes.MappingScene(1003)
    pts, f = [], 0
    while sum(p.shape[0] for p in pts) < 1_000_000:
        pts.append(sc.scan(f))
        f += 1
    pts = np.concatenate(pts)
    n = pts.shape[0]
    ids = np.arange(n, dtype=np.uint32)
    t = ikd.KD_TREE(0.5, 0.6, 0.5)
    t.Build(pts)
    rng = np.random.default_rng(3)
    live = np.ones(n, bool)

    # box queries centred on live points, sides log-uniform 1-20 m
    c = pts[rng.integers(0, n, 300)]
    side = np.exp(rng.uniform(0.0, np.log(20.0), (300, 3))).astype(np.float32)
    boxes = np.concatenate([c - side / 2, c + side / 2], axis=1).astype(np.float32)
    off, _, gi = t.box_search_batch(boxes)
    for i in range(0, 300, 3):
        want = ids[live & in_box(pts, boxes[i])]
        assert np.array_equal(np.sort(gi[off[i]:off[i + 1]]), np.sort(want)), f"box {i}"

    # sliding-window trims along the path, every 10 frames, with rebuilds
    for step, frame in enumerate(range(10, 60, 10)):
        slabs = _slabs(sc.pose(frame), 60.0 - 5.0 * step)
        cover = np.zeros(n, bool)
        for b in slabs:
            cover |= in_box(pts, b)
        want = int((live & cover).sum())
        assert t.Delete_Point_Boxes(slabs) == want, f"trim {step}"
        live &= ~cover
        assert t.validnum() == int(live.sum())
    stats = t.pool_stats()
    assert stats["rebuilds"] > 0

    xyz, gid = t.flatten()
    assert_same_multiset(xyz, pts[live])
    assert np.array_equal(np.sort(gid), ids[live])

    # radius queries centred on live points
    lp = np.nonzero(live)[0]
    cen = pts[lp[rng.integers(0, lp.size, 100)]]
    r = (0.2 + 2.8 * rng.random(100)).astype(np.float32)
    off, _, gi = t.radius_search_batch(cen, r)
    for i in range(100):
        inside, band = brute_radius(pts[live], cen[i], float(r[i]))
        got = set(gi[off[i]:off[i + 1]].tolist())
        lid = ids[live]
        must = set(lid[inside & ~band].tolist())
        may = set(lid[inside | band].tolist())
        assert must <= got <= may, f"radius {i}"

    # kNN after the trims
    q = pts[lp[rng.integers(0, lp.size, 300)]] + rng.normal(0, 0.3, (300, 3)).astype(np.float32)
    _, d, gi, cnt = t.knn(q, 5, 5.0)
    bd, bi, bc = brute_knn(pts[live], ids[live], q, 5, 5.0, tie="xyz")
    assert np.array_equal(cnt, bc)
    assert np.array_equal(d.view(np.uint32), bd.view(np.uint32))
    assert np.array_equal(gi, bi)
