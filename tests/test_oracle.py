"""Pins the CPU oracle (oracle/ikd_oracle.cpp) against checks that do not share
its code: numpy brute force, the half-open box rule, an independent flat-list
model of Add_Points' downsample rule, a recursive median definition of Build,
and tree invariants.  Each check is chosen so that a plausible slip in the
oracle (a dropped distance term, a closed box face, a wrong tie rule, a wrong
median index) fails it.
"""
import numpy as np
import pytest

from common import (INF, assert_same_multiset, brute_knn, brute_radius, check_dump_invariants, d2_all, grid_boxes,
                    grid_cloud, in_box, knn_equivalent, random_boxes, uniform_cloud)
from oracle import OracleTree


def _tree(points, ids=None, **kw):
    t = OracleTree(**kw)
    t.Build(points, ids)
    return t


# ------------------------------------------------------------------ kNN
@pytest.mark.parametrize("k,max_dist", [(1, INF), (5, INF), (5, 5.0), (20, 3.0), (10, 1.0)])
def test_knn_matches_bruteforce(k, max_dist):
    rng = np.random.default_rng(11 + k)
    pts = uniform_cloud(rng, 20000)
    q = uniform_cloud(rng, 600)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    t = _tree(pts, ids)
    xyz, d2, oid, cnt = t.knn(q, k, max_dist)
    bd, bid, bc = brute_knn(pts, ids, q, k, max_dist)
    assert knn_equivalent(d2, oid, cnt, bd, bid, bc, strict_ties=True) == []
    # returned coordinates are the stored points and d2 is their calc_dist
    for i in range(0, q.shape[0], 37):
        for j in range(cnt[i]):
            assert np.array_equal(xyz[i, j], pts[oid[i, j]])
            assert d2[i, j] == d2_all(pts[oid[i, j]][None], q[i])[0]
    # ascending output order (cpp:388-395)
    for i in range(q.shape[0]):
        c = cnt[i]
        assert np.all(np.diff(d2[i, :c]) >= 0)


def test_knn_ties_on_grid():
    """Grid points give many exact ties; the tie-insensitive parts must be exact."""
    rng = np.random.default_rng(5)
    pts = np.unique(grid_cloud(rng, 6000, extent=12), axis=0)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    q = grid_cloud(rng, 300, extent=12) + np.float32(0.125)
    t = _tree(pts, ids)
    _, d2, oid, cnt = t.knn(q, 8)
    bd, bid, bc = brute_knn(pts, ids, q, 8)
    assert knn_equivalent(d2, oid, cnt, bd, bid, bc) == []
    # d2 lists are always identical, even where the ids differ by a tie
    assert np.array_equal(np.sort(d2, axis=1), np.sort(bd, axis=1))


def test_knn_max_dist_is_a_double_comparison():
    """d <= max_dist*max_dist is evaluated in double (cpp:872,887): a point whose
    float d2 rounds to exactly max_dist^2 is kept, one ulp above is not."""
    q = np.zeros((1, 3), np.float32)
    r = np.float32(0.1)  # 0.1f squared in double is not a float
    m = float(r) * float(r)
    below = np.float32(np.sqrt(m))
    pts = np.array([[below, 0, 0], [np.nextafter(below, np.float32(1)), 0, 0], [5, 5, 5]], np.float32)
    t = _tree(pts)
    _, d2, _, cnt = t.knn(q, 3, float(r))
    bd, _, bc = brute_knn(pts, np.arange(3, dtype=np.uint32), q, 3, float(r))
    assert cnt[0] == bc[0]
    assert np.array_equal(d2[0, :cnt[0]], bd[0, :bc[0]])


def test_knn_rejects_nonpositive_k():
    t = _tree(uniform_cloud(np.random.default_rng(0), 100))
    with pytest.raises(ValueError):
        t.knn(np.zeros((1, 3), np.float32), 0)


# ------------------------------------------------------------- box search
@pytest.mark.parametrize("grid", [False, True])
def test_box_search_half_open(grid):
    rng = np.random.default_rng(21 + grid)
    if grid:
        pts = grid_cloud(rng, 8000, extent=20)
        boxes = grid_boxes(rng, 40, extent=20, smax=16)
    else:
        pts = uniform_cloud(rng, 8000)
        boxes = random_boxes(rng, 40)
    t = _tree(pts)
    for b in boxes:
        got, _ = t.Box_Search(b)
        assert_same_multiset(got, pts[in_box(pts, b)], "box search")


def test_radius_search_outside_band():
    rng = np.random.default_rng(31)
    pts = uniform_cloud(rng, 10000)
    t = _tree(pts, np.arange(10000, dtype=np.uint32))
    for i in range(60):
        q = uniform_cloud(rng, 1)[0]
        r = float(0.2 + 6.0 * rng.random())
        _, gid = t.Radius_Search(q, r)
        inside, band = brute_radius(pts, q, r)
        got = np.zeros(pts.shape[0], bool)
        got[gid] = True
        assert np.array_equal(got[~band], inside[~band])
        assert len(gid) == len(set(gid.tolist()))


# ---------------------------------------------------------------- deletes
@pytest.mark.parametrize("grid", [False, True])
def test_delete_point_boxes_count_and_live_set(grid):
    rng = np.random.default_rng(41 + grid)
    if grid:
        pts = grid_cloud(rng, 10000, extent=20)
        boxes = grid_boxes(rng, 10, extent=20, smax=20)
    else:
        pts = uniform_cloud(rng, 10000)
        boxes = random_boxes(rng, 10)
    t = _tree(pts)
    gone = np.zeros(pts.shape[0], bool)
    for b in boxes:
        gone |= in_box(pts, b)
    assert t.Delete_Point_Boxes(boxes) == int(gone.sum())
    assert t.validnum() == int((~gone).sum())
    live, _ = t.flatten()
    assert_same_multiset(live, pts[~gone], "live set after box delete")
    # queries see only live points
    q = uniform_cloud(rng, 200)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    _, d2, oid, cnt = t.knn(q, 5)
    bd, bid, bc = brute_knn(pts[~gone], ids[~gone], q, 5)
    assert knn_equivalent(d2, oid, cnt, bd, bid, bc, strict_ties=not grid) == []


def test_delete_points_distinct():
    rng = np.random.default_rng(51)
    pts = uniform_cloud(rng, 5000)
    t = _tree(pts)
    sel = rng.choice(5000, 700, replace=False)
    t.Delete_Points(pts[sel])
    keep = np.ones(5000, bool)
    keep[sel] = False
    assert t.validnum() == 4300
    assert_same_multiset(t.flatten()[0], pts[keep], "live set after Delete_Points")
    # deleting again, or deleting absent points, changes nothing
    t.Delete_Points(pts[sel[:50]])
    t.Delete_Points(uniform_cloud(rng, 50) + np.float32(200))
    assert t.validnum() == 4300


def test_delete_points_same_point_epsilon():
    """same_point uses |delta| < 1e-6 per axis (cpp:1368-1371)."""
    pts = np.array([[1.0, 2.0, 3.0], [10, 10, 10], [-4, 5, 6]], np.float32)
    t = _tree(pts)
    near = pts[0] + np.float32(5e-7)
    far = pts[2] + np.float32(3e-6)
    t.Delete_Points(np.stack([near, far]))
    assert_same_multiset(t.flatten()[0], pts[1:], "epsilon delete")


# ------------------------------------------------------------------- insert
def test_add_points_without_downsample_returns_zero():
    rng = np.random.default_rng(61)
    base = uniform_cloud(rng, 3000)
    extra = uniform_cloud(rng, 2000)
    t = _tree(base)
    assert t.Add_Points(extra, False) == 0  # cpp:470-488 never counts
    assert t.validnum() == 5000
    assert_same_multiset(t.flatten()[0], np.concatenate([base, extra]), "live set after insert")
    q = uniform_cloud(rng, 200)
    allp = np.concatenate([base, extra])
    ids = np.arange(5000, dtype=np.uint32)
    t2 = OracleTree()
    t2.Build(base, ids[:3000])
    t2.Add_Points(extra, False, ids[3000:])
    _, d2, oid, cnt = t2.knn(q, 5)
    bd, bid, bc = brute_knn(allp, ids, q, 5)
    assert knn_equivalent(d2, oid, cnt, bd, bid, bc, strict_ties=True) == []


def flat_downsample_model(live: list, batch: np.ndarray, ds: float):
    """Add_Points(downsample_on=true) restated on a flat list of live points
    (no tree), following cpp:422-448 point by point."""
    added = 0
    ds32 = np.float32(ds)
    for p in batch:
        lo = np.floor(p / ds32).astype(np.float32) * ds32
        hi = (lo + ds32).astype(np.float32)
        mid = np.array([float(lo[a]) + float(hi[a] - lo[a]) / 2.0 for a in range(3)], np.float32)
        arr = np.array(live, np.float32).reshape(-1, 3)
        inb = np.nonzero(np.all((arr >= lo) & (arr < hi), axis=1))[0] if arr.size else np.zeros(0, int)
        best = p
        best_d = d2_all(p[None], mid)[0]
        if inb.size:
            dd = d2_all(arr[inb], mid)
            j = int(np.argmin(dd))
            if dd[j] < best_d:
                best = arr[inb[j]]
        same = bool(np.all(np.abs(p - best) < 1e-6))
        if inb.size > 1 or same:
            if inb.size:
                keep = np.ones(arr.shape[0], bool)
                keep[inb] = False
                live = [tuple(r) for r in arr[keep]]
            live.append(tuple(best))
            added += 1
    return live, added


@pytest.mark.parametrize("ds", [0.5, 1.0, 0.2])
def test_add_points_downsample_matches_flat_model(ds):
    rng = np.random.default_rng(71)
    base = uniform_cloud(rng, 1500, -6, 6)
    t = OracleTree(box_length=ds)
    t.Build(base)
    live = [tuple(r) for r in base]
    total_o = total_m = 0
    for it in range(4):
        batch = uniform_cloud(rng, 300, -7, 7)
        # a few exact duplicates and near-duplicates exercise same_point
        batch[::50] = base[rng.integers(0, 1500, batch[::50].shape[0])]
        total_o += t.Add_Points(batch, True)
        live, a = flat_downsample_model(live, batch, ds)
        total_m += a
        assert total_o == total_m
        assert t.validnum() == len(live)
        assert_same_multiset(t.flatten()[0], np.array(live, np.float32), f"downsample live set (iter {it})")


# -------------------------------------------------------------------- build
def _ref_build(rows):
    """BuildTree's definition (cpp:573-622): axis of longest extent (ties to the
    lower axis), the element of rank (n-1)>>1 on that axis is the node."""
    out = []

    def rec(sub):
        if sub.shape[0] == 0:
            return
        ext = sub[:, :3].max(axis=0) - sub[:, :3].min(axis=0)
        axis = 0
        for a in (1, 2):
            if ext[a] > ext[axis]:
                axis = a
        s = sub[np.argsort(sub[:, axis], kind="stable")]
        mid = (s.shape[0] - 1) >> 1
        out.append((int(s[mid, 3]), axis))
        rec(s[:mid])
        rec(s[mid + 1:])

    rec(rows)
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 301])
def test_build_structure_matches_median_definition(n):
    rng = np.random.default_rng(80 + n)
    pts = uniform_cloud(rng, n)
    ids = np.arange(n, dtype=np.uint32)
    t = _tree(pts, ids)
    dump = t.dump()
    got = [(int(r["id"]), int(r["axis"])) for r in dump]
    rows = np.concatenate([pts.astype(np.float64), ids[:, None].astype(np.float64)], axis=1)
    assert got == _ref_build(rows)
    assert t.size() == n and t.validnum() == n


# --------------------------------------------------------------- invariants
def _check_invariants(t: OracleTree, tight: bool = True):
    d = t.dump()
    check_dump_invariants(d, tight)
    return d


def test_invariants_after_mixed_operations():
    rng = np.random.default_rng(91)
    t = OracleTree(0.5, 0.6, 0.5)
    t.Build(uniform_cloud(rng, 4000, -20, 20))
    for it in range(6):
        t.Add_Points(uniform_cloud(rng, 400, -22, 22), it % 2 == 0)
        _check_invariants(t)
        t.Delete_Point_Boxes(random_boxes(rng, 3, -20, 20, 2, 12))
        _check_invariants(t)
    live = t.flatten()[0]
    t.Delete_Points(live[::7])
    _check_invariants(t, tight=False)  # Delete_by_point leaves the node's own range stale
    assert t.validnum() == live.shape[0] - live[::7].shape[0]


def test_rebuild_collapses_deleted_nodes():
    """Deleting most points triggers the alpha_del rebuild (cpp:1089-1107):
    size() drops towards validnum() while the live set is unchanged."""
    rng = np.random.default_rng(101)
    pts = uniform_cloud(rng, 3000)
    t = _tree(pts)
    sel = rng.choice(3000, 2000, replace=False)
    t.Delete_Points(pts[sel])
    assert t.validnum() == 1000
    assert t.size() < 3000
    keep = np.ones(3000, bool)
    keep[sel] = False
    assert_same_multiset(t.flatten()[0], pts[keep], "live set after rebuilds")


def test_acquire_removed_points_records_rebuild_drops():
    """acquire_removed_points (cpp:558-571): a box delete of most points rebuilds
    the root, which drops and records exactly the points in the box; points the
    downsample deleted are never recorded; recorded + deleted still in the tree
    = deleted by the caller; the record is cleared by the call."""
    rng = np.random.default_rng(93)
    pts = uniform_cloud(rng, 8000, 0, 20)
    ids = np.arange(8000, dtype=np.uint32)
    o = OracleTree(0.5, 0.6, 0.5)
    o.Build(pts, ids)
    box = np.array([[-1, -1, -1, 13, 21, 21]], np.float32)
    n = o.Delete_Point_Boxes(box)
    x, i = o.acquire_removed_points()
    inside = in_box(pts, box[0])
    assert n == len(i) == int(inside.sum())
    assert sorted(i.tolist()) == sorted(ids[inside].tolist())
    assert np.array_equal(x[np.argsort(i)], pts[np.sort(i)])
    assert len(o.acquire_removed_points()[1]) == 0
    # downsample deletions: voxel centres replace the old points of their voxels
    live, live_ids = o.flatten()
    cc = (np.floor(live[:200] / 0.5) * 0.5 + 0.25).astype(np.float32)
    o.Add_Points(cc, True, np.arange(10000, 10200, dtype=np.uint32))
    deleted = 0
    recorded = set()
    for it in range(4):
        live, live_ids = o.flatten()
        sel = rng.choice(live.shape[0], live.shape[0] // 3, replace=False)
        v0 = o.validnum()
        o.Delete_Points(live[sel])
        deleted += v0 - o.validnum()
        _, ri = o.acquire_removed_points()
        assert not (set(ri.tolist()) & recorded)
        recorded |= set(ri.tolist())
        assert recorded <= set(range(8000)) | set(range(10000, 10200))
    ds_size = o.size() - o.validnum()  # deleted nodes still in the tree, downsample ones included
    assert len(recorded) <= deleted <= len(recorded) + ds_size


# ------------------------------------------------------------ near ties
NEAR_TIE = np.array([[3e-6, 0, 0], [-7e-6, 0, 0], [0, 8e-6, 0]], np.float32)  # P1, P2, P3


def test_knn_near_tie_comparator():
    """PointType_CMP (ikd_Tree.h:99-102) orders distances closer than 1e-10 by
    x.  Around the origin d2(P1) = 9e-12 < d2(P2) = 4.9e-11 < d2(P3) = 6.4e-11,
    all within 1e-10, while x orders them P2 < P3 < P1.  BuildTree puts P3
    (median x) at the root, P2 left and P1 right.  Search (cpp:868-1013) for
    k = 2 takes P3, then the nearer child P1; the heap's top is then P1 (the
    largest x, not the largest distance), so the left child's box distance
    4.9e-11 is not below top.dist = 9e-12 and P2 is never reached.  The
    result is {P3, P1}, popped top-first into the front: P3 then P1."""
    t = OracleTree()
    t.Build(NEAR_TIE)
    d = t.dump()
    assert [int(r["id"]) for r in d] == [2, 1, 0]  # pre-order: P3, P2, P1
    xyz, d2, ids = t.Nearest_Search([0, 0, 0], 2)
    assert ids.tolist() == [2, 0]
    q = np.zeros(3, np.float32)
    assert d2.view(np.uint32).tolist() == [d2_all(NEAR_TIE[2:3], q)[0].view(np.uint32),
                                           d2_all(NEAR_TIE[0:1], q)[0].view(np.uint32)]
    # the exact 2 nearest are P1 and P2; P2 and P3 lie within the comparator's
    # 1e-10 window of each other, which knn_equivalent treats as a tie
    bd, bi, bc = brute_knn(NEAR_TIE, np.arange(3, dtype=np.uint32), q[None], 2)
    assert bi[0].tolist() == [0, 1]
    _, od, oi, oc = t.knn(q[None], 2)
    assert knn_equivalent(od, oi, oc, bd, bi, bc, strict_ties=True) == [0]
    assert knn_equivalent(od, oi, oc, bd, bi, bc) == []
