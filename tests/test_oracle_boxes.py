"""The oracle's Add_Point_Boxes (Add_by_range, ikd-Tree/ikd_Tree.cpp:762-815)
pinned without the oracle's own formula:

- hand-built trees whose outcome follows from the reference's rules on paper
  (a covered node brings back its whole subtree, points outside the box
  included; a box that misses every stored range brings back nothing);
- a closed form: with rebuilds off, a box that covers the root's range gives
  back every point ever inserted except those the downsample deleted, which
  an independent flat model of Add_Points' downsample rule lists;
- set laws that hold with rebuilds on: restoring never deletes, and every
  point it brings back was deleted, not by the downsample, and never left the
  tree through a rebuild.
"""
import numpy as np
import pytest

from common import assert_same_multiset, random_boxes, uniform_cloud
from oracle import OracleTree
from test_oracle import flat_downsample_model

NO_REBUILD = (1.0, 1.0)  # delete/balance parameters no Criterion_Check can exceed (cpp:1089-1107)


def _line(n):
    return np.stack([np.arange(n, dtype=np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)], axis=1)


def _box(x0, x1):
    return np.array([[x0, -1, -1, x1, 1, 1]], np.float32)


def test_covered_root_brings_back_points_outside_the_box():
    # x = 0..6: root 3, left subtree {0,1,2}, right subtree {4,5,6} (7 nodes,
    # below Minimal_Unbalanced_Tree_Size, so nothing is rebuilt)
    t = OracleTree()
    t.Build(_line(7))
    assert t.Delete_Point_Boxes(_box(3.5, 7)) == 3  # covers the right subtree
    assert np.array_equal(t.tree_range(), np.array([0, 0, 0, 3, 0, 0], np.float32))
    # the box covers the root's stored range [0,3]: the root's labels are reset
    # and Push_Down hands the reset to the deleted right subtree too
    t.Add_Point_Boxes(_box(-1, 3.5))
    assert t.validnum() == 7
    assert_same_multiset(t.flatten()[0], _line(7), "restored set")


def test_box_missing_every_stored_range_brings_back_nothing():
    t = OracleTree()
    t.Build(_line(7))
    t.Delete_Point_Boxes(_box(3.5, 7))
    # the deleted points lie in this box, but the root's range [0,3] does not
    # reach it, so the walk stops at the root
    t.Add_Point_Boxes(_box(3.5, 7))
    assert t.validnum() == 4
    assert_same_multiset(t.flatten()[0], _line(4), "nothing restored")


def test_cut_node_gets_back_only_its_own_point():
    t = OracleTree()
    t.Build(_line(7))
    # delete 1 and 5 one by one (internal nodes: their subtrees stay live)
    t.Delete_Points(np.array([[1, 0, 0], [5, 0, 0]], np.float32))
    assert t.validnum() == 5
    # [0.5, 1.5) cuts the root and the left subtree's root (range [0,2]) and
    # holds point 1 only
    t.Add_Point_Boxes(_box(0.5, 1.5))
    assert t.validnum() == 6
    assert_same_multiset(t.flatten()[0], _line(7)[[0, 1, 2, 3, 4, 6]], "point 1 back")


def test_downsample_deleted_points_stay_deleted_closed_form():
    rng = np.random.default_rng(91)
    base = uniform_cloud(rng, 2000, -6, 6)
    t = OracleTree(*NO_REBUILD, 0.5)
    t.Build(base)
    live = [tuple(r) for r in base]
    ever = [tuple(r) for r in base]
    for _ in range(3):
        batch = uniform_cloud(rng, 400, -7, 7)
        t.Add_Points(batch, True)
        before = set(live)
        live, _ = flat_downsample_model(live, batch, 0.5)
        ever += [p for p in live if p not in before]
    ds_deleted = set(ever) - set(live)
    # box deletes (not downsample) of a few regions, then one box over everything
    t.Delete_Point_Boxes(random_boxes(rng, 6, -6, 6, 2, 6))
    assert t.validnum() < len(live)
    t.Add_Point_Boxes(np.array([[-1e6, -1e6, -1e6, 1e6, 1e6, 1e6]], np.float32))
    expect = np.array([p for p in ever if p not in ds_deleted], np.float32)
    assert t.validnum() == expect.shape[0]
    assert_same_multiset(t.flatten()[0], expect, "everything but the downsample's deletions")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_restore_laws_with_rebuilds(seed):
    rng = np.random.default_rng(300 + seed)
    pts = uniform_cloud(rng, 6000)
    ids = np.arange(6000, dtype=np.uint32)
    t = OracleTree()
    t.Build(pts, ids)
    removed = set()
    for _ in range(4):
        t.Delete_Point_Boxes(random_boxes(rng, 3, smin=8, smax=30))
        removed |= set(t.acquire_removed_points()[1].tolist())
        before = set(t.flatten()[1].tolist())
        t.Add_Point_Boxes(random_boxes(rng, 2, smin=10, smax=40))
        removed |= set(t.acquire_removed_points()[1].tolist())
        after = set(t.flatten()[1].tolist())
        assert before <= after, "Add_Point_Boxes deleted points"
        back = after - before
        assert not (back & removed), "a point a rebuild removed came back"
        assert t.validnum() == len(after)
