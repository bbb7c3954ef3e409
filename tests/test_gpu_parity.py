"""GPU parity: the HIP path through the C ABI against numpy brute force and the
CPU oracle, on the same seeded inputs.

Bars (what "equal" means here):
- kNN: the device returns the k smallest live points under (d2, x, y, z, id);
  against brute force with that order, squared distances, ids and counts are
  bit-exact.  Against the oracle (the reference's own heap and traversal) the
  squared distances are bit-exact and the ids equal except for entries tied
  at the k-th distance (the reference keeps whichever it met first).
- box search, box delete, Add_Points, Delete_Points, flatten: exact multisets
  of points and exact counts.
- radius search: exact outside a 4-ulp band around the sphere (the
  reference's sqrt-based subtree rule can resolve band points either way).
"""
import importlib

import numpy as np
import pytest

from common import (INF, assert_same_multiset, brute_knn, brute_radius, grid_boxes, grid_cloud, in_box,
                    knn_equivalent, random_boxes, uniform_cloud)
from oracle import OracleTree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ikd():
    return importlib.import_module("ikd-tree_amd")


def _pair(ikd, pts, ids=None, params=(0.5, 0.6, 0.2)):
    g = ikd.KD_TREE(*params)
    o = OracleTree(*params)
    g.Build(pts, ids)
    o.Build(pts, ids)
    return g, o


def _check_knn(g, o, pts_live, ids_live, q, k, max_dist, strict_vs_oracle=True):
    gx, gd, gi, gc = g.knn(q, k, max_dist)
    bd, bi, bc = brute_knn(pts_live, ids_live, q, k, max_dist, tie="xyz")
    assert np.array_equal(gc, bc), "counts vs brute force"
    for i in range(q.shape[0]):
        c = gc[i]
        assert np.array_equal(gd[i, :c].view(np.uint32), bd[i, :c].view(np.uint32)), f"d2 q{i}"
        assert np.array_equal(gi[i, :c], bi[i, :c]), f"ids q{i}"
    # returned coordinates are the stored points
    lut = {int(i): p for i, p in zip(ids_live, pts_live)}
    for i in range(0, q.shape[0], 17):
        for j in range(gc[i]):
            assert np.array_equal(gx[i, j], lut[int(gi[i, j])])
    od = o.knn(q, k, max_dist)
    assert knn_equivalent(gd, gi, gc, od[1], od[2], od[3], strict_ties=strict_vs_oracle) == []


@pytest.mark.parametrize("k,max_dist", [(1, INF), (5, INF), (5, 5.0), (8, 2.0), (20, INF), (64, 3.0)])
def test_build_knn_uniform(ikd, k, max_dist):
    rng = np.random.default_rng(100 + k)
    pts = uniform_cloud(rng, 60000)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    q = uniform_cloud(rng, 3000)
    g, o = _pair(ikd, pts, ids)
    assert g.validnum() == 60000 and g.size() == 60000
    _check_knn(g, o, pts, ids, q, k, max_dist)
    assert np.array_equal(g.tree_range(), o.tree_range())


def _same_structure(gd, od):
    assert gd.shape == od.shape, f"node counts {gd.shape[0]} vs {od.shape[0]}"
    for f in ("id", "axis", "tree_size", "invalid", "left", "right"):
        assert np.array_equal(gd[f], od[f]), f
    for f in ("x", "y", "z"):
        assert np.array_equal(gd[f].view(np.uint32), od[f].view(np.uint32)), f
    assert np.array_equal(gd["range"].view(np.uint32), od["range"].view(np.uint32)), "range"
    assert np.array_equal(gd["flags"] & 3, od["flags"] & 3), "flags"


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 1000, 1024, 1025, 5000])
@pytest.mark.parametrize("small", [1, 0])
def test_build_structure_matches_oracle(ikd, n, small):
    """Tie-free keys: the device build (one-workgroup build for <= 1024 points,
    level-synchronous otherwise) is node for node the reference's BuildTree."""
    rng = np.random.default_rng(7 + n)
    pts = uniform_cloud(rng, n)
    g = ikd.KD_TREE()
    g.set_tuning(ikd.IKD_TUNE_SMALL_BUILD, small)
    g.Build(pts)
    o = OracleTree()
    o.Build(pts)
    _same_structure(g.dump_nodes(), o.dump())


@pytest.mark.parametrize("n", [700, 1500, 5000, 40000])
def test_build_paths_identical_with_ties(ikd, n):
    """Grid points (many equal keys on every axis): the level-synchronous build
    and the hybrid build (levels, then one workgroup per small segment) place
    equal keys the same way, so the two trees are node for node the same."""
    rng = np.random.default_rng(70 + n)
    pts = grid_cloud(rng, n, step=0.5, extent=12)
    dumps = []
    for small in (0, 1):
        g = ikd.KD_TREE()
        g.set_tuning(ikd.IKD_TUNE_SMALL_BUILD, small)
        g.Build(pts)
        dumps.append(g.dump_nodes())
    _same_structure(dumps[0], dumps[1])


def test_morton_order_does_not_change_results(ikd):
    rng = np.random.default_rng(7)
    pts = uniform_cloud(rng, 5000)
    g, o = _pair(ikd, pts)
    g.set_tuning(ikd.IKD_TUNE_QUERY_ORDER, 1)
    q = uniform_cloud(rng, 6000)  # >= 4096 queries takes the Morton-sorted path
    _, d1, i1, c1 = g.knn(q, 5)
    _, d2, i2, c2 = g.knn(q[:100], 5)
    assert np.array_equal(d1[:100], d2) and np.array_equal(i1[:100], i2)


@pytest.mark.parametrize("k,max_dist", [(1, INF), (5, 5.0), (10, INF), (20, 2.0)])
def test_knn_lanes_per_query_invariance(ikd, k, max_dist):
    """Every lanes-per-query choice (team traversal) returns the same bits."""
    rng = np.random.default_rng(300 + k)
    pts = uniform_cloud(rng, 80000)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    q = uniform_cloud(rng, 2500)
    g, o = _pair(ikd, pts, ids)
    ref = None
    for lanes in (1, 0, 2, 4, 8, 16):
        g.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, lanes)
        out = g.knn(q, k, max_dist)
        if ref is None:
            ref = out
            bd, bi, bc = brute_knn(pts, ids, q, k, max_dist, tie="xyz")
            assert np.array_equal(out[3], bc)
            for i in range(q.shape[0]):
                c = bc[i]
                assert np.array_equal(out[1][i, :c].view(np.uint32), bd[i, :c].view(np.uint32))
                assert np.array_equal(out[2][i, :c], bi[i, :c])
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                      b.view(np.uint32) if b.dtype == np.float32 else b), f"lanes={lanes}"


@pytest.mark.parametrize("k,max_dist", [(5, INF), (8, 2.0)])
def test_knn_lds_staging_invariance(ikd, k, max_dist):
    """Staging the tree's top levels in LDS (IKD_TUNE_KNN_STAGE 1-12, capped
    by the LDS budget) returns the same bits as the plain one-lane walk, on a
    tree with deleted points and deleted subtrees."""
    rng = np.random.default_rng(340 + k)
    pts = uniform_cloud(rng, 60000)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    g, o = _pair(ikd, pts, ids)
    g.Delete_Point_Boxes(random_boxes(rng, 6, smin=5, smax=25))
    g.Delete_Points(pts[rng.choice(pts.shape[0], 3000, replace=False)])
    q = uniform_cloud(rng, 3000)
    g.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, 1)
    ref = g.knn(q, k, max_dist)
    for levels in (1, 3, 6, 9, 12):
        g.set_tuning(ikd.IKD_TUNE_KNN_STAGE, levels)
        out = g.knn(q, k, max_dist)
        for a_, b_ in zip(out, ref):
            assert np.array_equal(a_.view(np.uint32) if a_.dtype == np.float32 else a_,
                                  b_.view(np.uint32) if b_.dtype == np.float32 else b_), f"levels={levels}"
    live, lid = g.flatten()
    bd, bi, bc = brute_knn(live, lid, q, k, max_dist, tie="xyz")
    assert np.array_equal(ref[3], bc)
    for i in range(q.shape[0]):
        assert np.array_equal(ref[1][i, :bc[i]].view(np.uint32), bd[i, :bc[i]].view(np.uint32))


@pytest.mark.parametrize("k,max_dist", [(1, INF), (5, 5.0), (20, INF)])
def test_knn_compact_stack_invariance(ikd, k, max_dist):
    """The 4-byte LDS stack entries (power-of-two distance bound) return the
    same bits as the 6-byte ones, for one lane per query, teams of lanes and
    persistent lanes, on a tree with deleted points and subtrees and with
    exact distance ties."""
    rng = np.random.default_rng(360 + k)
    pts = np.concatenate([uniform_cloud(rng, 50000), grid_cloud(rng, 20000, extent=16)])
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    g, o = _pair(ikd, pts, ids)
    g.Delete_Point_Boxes(random_boxes(rng, 6, smin=5, smax=25))
    g.Delete_Points(pts[rng.choice(pts.shape[0], 3000, replace=False)])
    q = np.concatenate([uniform_cloud(rng, 2500), grid_cloud(rng, 500, extent=16) + np.float32(0.125)])
    out = []
    for lanes in (1, 4, 0):  # one lane, a team of 4, persistent lanes
        g.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, lanes)
        for compact in (0, 1):
            g.set_tuning(ikd.IKD_TUNE_KNN_COMPACT, compact)
            out.append(g.knn(q, k, max_dist))
    for o in out[1:]:
        for a_, b_ in zip(o, out[0]):
            assert np.array_equal(a_.view(np.uint32) if a_.dtype == np.float32 else a_,
                                  b_.view(np.uint32) if b_.dtype == np.float32 else b_)
    live, lid = g.flatten()
    bd, bi, bc = brute_knn(live, lid, q, k, max_dist, tie="xyz")
    assert np.array_equal(out[1][3], bc)
    for i in range(q.shape[0]):
        assert np.array_equal(out[1][1][i, :bc[i]].view(np.uint32), bd[i, :bc[i]].view(np.uint32))


def test_box_search_order_is_preorder(ikd):
    """A box holding the whole tree returns the live points in flatten's
    pre-order (own point, left subtree, right subtree), through the walk and
    the covered-subtree tasks alike; boxes of every size return exact sets."""
    rng = np.random.default_rng(350)
    pts = uniform_cloud(rng, 50000)
    g, o = _pair(ikd, pts)
    g.Delete_Point_Boxes(random_boxes(rng, 4, smin=5, smax=30))
    g.Delete_Points(pts[rng.choice(pts.shape[0], 2000, replace=False)])
    live, lid = g.flatten()
    big = np.array([[-1e6, -1e6, -1e6, 1e6, 1e6, 1e6]], np.float32)
    off, xyz, got_ids = g.box_search_batch(big)
    assert np.array_equal(got_ids[:off[1]], lid) and np.array_equal(xyz[:off[1]], live)
    boxes = np.concatenate([random_boxes(rng, 60, smin=1, smax=10), random_boxes(rng, 60, smin=30, smax=90)])
    off, xyz, _ = g.box_search_batch(boxes)
    for b in range(boxes.shape[0]):
        assert_same_multiset(xyz[off[b]:off[b + 1]], live[in_box(live, boxes[b])], f"box {b}")


def test_knn_grid_ties(ikd):
    rng = np.random.default_rng(9)
    pts = grid_cloud(rng, 20000, extent=16)  # duplicates and exact distance ties
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    q = grid_cloud(rng, 1500, extent=16) + np.float32(0.125)
    g, o = _pair(ikd, pts, ids)
    _check_knn(g, o, pts, ids, q, 10, INF, strict_vs_oracle=False)


def test_empty_and_tiny_trees(ikd):
    g = ikd.KD_TREE()
    q = np.zeros((5, 3), np.float32)
    _, d, i, c = g.knn(q, 3)
    assert np.all(c == 0) and np.all(np.isinf(d))
    assert g.validnum() == 0 and g.size() == 0
    g.Build(np.array([[1, 2, 3]], np.float32))
    _, d, i, c = g.knn(q, 3)
    assert np.all(c == 1) and np.all(d[:, 0] == np.float32(14.0))
    g.Build(np.zeros((0, 3), np.float32))
    assert g.validnum() == 0


def test_argument_errors(ikd):
    g = ikd.KD_TREE()
    g.Build(uniform_cloud(np.random.default_rng(1), 100))
    with pytest.raises(ikd.IkdError) as e:
        g.knn(np.zeros((1, 3), np.float32), 0)
    assert e.value.code == ikd.IKD_ERR_ARG
    with pytest.raises(ikd.IkdError) as e:
        g.knn(np.zeros((1, 3), np.float32), 65)
    assert e.value.code == ikd.IKD_ERR_UNSUPPORTED


def test_downsample_out_of_range_voxel_changes_nothing(ikd):
    """A downsampled batch holding a point whose voxel index floor(p / box)
    is beyond +-2^20 is refused (IKD_ERR_UNSUPPORTED) before any of its
    deletions lands: the tree keeps every point and answers as before."""
    rng = np.random.default_rng(5)
    g = ikd.KD_TREE(0.5, 0.6, 0.5)
    base = uniform_cloud(rng, 5000, -5, 5)
    g.Build(base)
    q = uniform_cloud(rng, 500, -6, 6)
    _, d0, i0, c0 = g.knn(q, 5)
    before, bid = g.flatten()
    batch = uniform_cloud(rng, 2000, -5, 5)  # its voxels hold base points the fold would delete
    batch[17] = [1e9, 0.25, 0.25]
    with pytest.raises(ikd.IkdError) as e:
        g.Add_Points(batch, True)
    assert e.value.code == ikd.IKD_ERR_UNSUPPORTED
    after, aid = g.flatten()
    assert g.validnum() == 5000
    assert_same_multiset(before, after)
    assert np.array_equal(np.sort(bid), np.sort(aid))
    _, d1, i1, c1 = g.knn(q, 5)
    assert np.array_equal(c0, c1) and np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
    assert np.array_equal(i0, i1)


@pytest.mark.parametrize("grid", [False, True])
def test_box_search(ikd, grid):
    rng = np.random.default_rng(20 + grid)
    if grid:
        pts = grid_cloud(rng, 30000, extent=20)
        boxes = grid_boxes(rng, 200, extent=20, smax=16)
    else:
        pts = uniform_cloud(rng, 30000)
        boxes = random_boxes(rng, 200, smin=1, smax=20)
    g, o = _pair(ikd, pts)
    off, xyz, _ = g.box_search_batch(boxes)
    for b in range(boxes.shape[0]):
        got = xyz[off[b]:off[b + 1]]
        assert_same_multiset(got, pts[in_box(pts, boxes[b])], f"box {b}")
        if b % 20 == 0:
            assert_same_multiset(got, o.Box_Search(boxes[b])[0], f"box {b} vs oracle")


def test_radius_search(ikd):
    rng = np.random.default_rng(30)
    pts = uniform_cloud(rng, 30000)
    ids = np.arange(30000, dtype=np.uint32)
    g, o = _pair(ikd, pts, ids)
    c = uniform_cloud(rng, 300)
    r = (0.2 + 6.0 * rng.random(300)).astype(np.float32)
    off, _, gid = g.radius_search_batch(c, r)
    for i in range(300):
        got = np.zeros(30000, bool)
        got[gid[off[i]:off[i + 1]]] = True
        inside, band = brute_radius(pts, c[i], float(r[i]))
        assert np.array_equal(got[~band], inside[~band]), f"radius query {i}"
        if i % 30 == 0:
            _, oid = o.Radius_Search(c[i], float(r[i]))
            og = np.zeros(30000, bool)
            og[oid] = True
            assert np.array_equal(got[~band], og[~band])


@pytest.mark.parametrize("grid", [False, True])
def test_delete_point_boxes(ikd, grid):
    rng = np.random.default_rng(40 + grid)
    if grid:
        pts = grid_cloud(rng, 40000, extent=20)
        boxes = grid_boxes(rng, 10, extent=20, smax=24)
    else:
        pts = uniform_cloud(rng, 40000)
        boxes = random_boxes(rng, 10)
    ids = np.arange(pts.shape[0], dtype=np.uint32)
    g, o = _pair(ikd, pts, ids)
    gone = np.zeros(pts.shape[0], bool)
    for b in boxes:
        gone |= in_box(pts, b)
    n_g = g.Delete_Point_Boxes(boxes)
    n_o = o.Delete_Point_Boxes(boxes)
    assert n_g == n_o == int(gone.sum())
    assert g.validnum() == o.validnum() == int((~gone).sum())
    assert_same_multiset(g.flatten()[0], pts[~gone], "live set after box delete")
    q = uniform_cloud(rng, 1500)
    _check_knn(g, o, pts[~gone], ids[~gone], q, 5, INF, strict_vs_oracle=not grid)
    # deleting again removes nothing
    assert g.Delete_Point_Boxes(boxes) == 0


def test_delete_everything_then_insert(ikd):
    rng = np.random.default_rng(45)
    pts = uniform_cloud(rng, 5000)
    g, o = _pair(ikd, pts)
    box = np.array([[-60, -60, -60, 60, 60, 60]], np.float32)
    assert g.Delete_Point_Boxes(box) == 5000
    assert g.validnum() == 0
    extra = uniform_cloud(rng, 300)
    assert g.Add_Points(extra, False) == 0
    o.Delete_Point_Boxes(box)
    o.Add_Points(extra, False)
    assert g.validnum() == 300
    assert_same_multiset(g.flatten()[0], extra, "reinsert after full delete")


def test_add_points_without_downsample(ikd):
    rng = np.random.default_rng(50)
    base = uniform_cloud(rng, 20000)
    ids = np.arange(40000, dtype=np.uint32)
    g, o = _pair(ikd, base, ids[:20000])
    allp = [base]
    for it in range(4):
        extra = uniform_cloud(rng, 5000, -60, 60)
        assert g.Add_Points(extra, False, ids[20000 + 5000 * it:20000 + 5000 * (it + 1)]) == 0
        o.Add_Points(extra, False, ids[20000 + 5000 * it:20000 + 5000 * (it + 1)])
        allp.append(extra)
    allp = np.concatenate(allp)
    assert g.validnum() == 40000
    assert_same_multiset(g.flatten()[0], allp, "live set after inserts")
    q = uniform_cloud(rng, 1500, -60, 60)
    _check_knn(g, o, allp, ids, q, 5, INF)


@pytest.mark.parametrize("ds", [0.5, 1.0, 0.2])
def test_add_points_downsample_vs_oracle(ikd, ds):
    rng = np.random.default_rng(60)
    base = uniform_cloud(rng, 8000, -10, 10)
    g, o = _pair(ikd, base, params=(0.5, 0.6, ds))
    for it in range(5):
        batch = uniform_cloud(rng, 3000, -11, 11)
        batch[::40] = base[rng.integers(0, 8000, batch[::40].shape[0])]  # exact duplicates
        a_g = g.Add_Points(batch, True)
        a_o = o.Add_Points(batch, True)
        assert a_g == a_o, f"Add_Points return, batch {it}"
        assert g.validnum() == o.validnum()
        assert_same_multiset(g.flatten()[0], o.flatten()[0], f"downsampled live set, batch {it}")


@pytest.mark.parametrize("ds", [0.5, 0.3])
def test_downsample_wide_batches_vs_oracle(ikd, ds):
    """Batches over a map 1.4 km wide (thousands of voxels per axis) and
    narrow ones after them, with exact and near duplicates of map points:
    every batch matches the oracle (0.3 m: the conflict check's neighbour
    lookups)."""
    rng = np.random.default_rng(62)
    base = uniform_cloud(rng, 6000, -700, 700)
    g, o = _pair(ikd, base, params=(0.5, 0.6, ds))
    for it, (lo, hi) in enumerate([(-700, 700), (-60, 60), (-700, 700), (-3, 3)]):
        batch = uniform_cloud(rng, 3000, lo, hi)
        batch[::10] = base[rng.integers(0, 6000, batch[::10].shape[0])]  # exact duplicates of map points
        batch[1::10] = batch[::10] + np.float32(0.01)  # and near ones in the same voxels
        a_g = g.Add_Points(batch, True)
        a_o = o.Add_Points(batch, True)
        assert a_g == a_o, f"Add_Points return, batch {it}"
        assert g.validnum() == o.validnum()
        assert_same_multiset(g.flatten()[0], o.flatten()[0], f"downsampled live set, batch {it}")


def _dense_voxel_cloud(rng):
    """Thousands of points in a few 0.5 m voxels, exact duplicates, and pairs of
    distinct points at the same distance from their voxel's centre (the
    reference keeps the first of them in Search_by_range's pre-order)."""
    dense = rng.uniform(0.0, 2.0, (12000, 3)).astype(np.float32)
    c = (np.floor(rng.uniform(4, 8, (300, 3))) * 0.5 + 0.25).astype(np.float32)  # voxel centres beside the cube
    e = np.float32(0.125)
    tie_a = c + np.array([e, 0, 0], np.float32)
    tie_b = c - np.array([e, 0, 0], np.float32)
    tie_c = c + np.array([0, 0, e], np.float32)
    return np.concatenate([dense, tie_a, tie_b, tie_c, dense[:500]])


def _voxel_signature(pts, ds):
    """(voxel index, squared distance to the voxel centre) of every point, sorted:
    equal for two live sets that keep, in every voxel, points at the same
    distances from its centre (ties between distinct equidistant points aside)."""
    v = np.floor(pts / np.float32(ds)).astype(np.int64)
    c = v.astype(np.float64) * ds + ds / 2
    d = ((pts.astype(np.float64) - c) ** 2).sum(axis=1)
    return sorted(zip(map(tuple, v.tolist()), d.tolist()))


def test_downsample_dense_voxels_vs_oracle(ikd):
    """Dense voxels with exact distance ties.  Which of several equidistant
    points a voxel keeps follows the pre-order of the tree at the start of the
    batch; the reference rebuilds inside its per-point loop, which can reorder
    the tied points first, so the comparison is per voxel and distance."""
    rng = np.random.default_rng(61)
    base = _dense_voxel_cloud(rng)
    g, o = _pair(ikd, base, params=(0.5, 0.6, 0.5))
    for it in range(4):
        batch = _dense_batch(rng, it)
        assert g.Add_Points(batch, True) == o.Add_Points(batch, True), f"Add_Points return, batch {it}"
        assert g.validnum() == o.validnum()
        assert _voxel_signature(g.flatten()[0], 0.5) == _voxel_signature(o.flatten()[0], 0.5), f"batch {it}"


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_downsample_fold_lanes_identical(ikd, lanes):
    """Teams of lanes per voxel keep exactly the points (and ids) one lane keeps,
    equidistant ties included."""
    rng = np.random.default_rng(62)
    base = _dense_voxel_cloud(rng)
    out = []
    for ln in (1, lanes):
        g = ikd.KD_TREE(0.5, 0.6, 0.5)
        g.Build(base)
        g.set_tuning(ikd.IKD_TUNE_FOLD_LANES, ln)
        r = np.random.default_rng(63)
        # one call after a Build: the tree the fold walks is the same in both
        added = g.Add_Points(np.concatenate([_dense_batch(r, 0), _dense_batch(r, 1)]), True)
        x, ids = g.flatten()
        order = np.lexsort((ids, x[:, 2], x[:, 1], x[:, 0]))
        out.append((added, x[order], ids[order]))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def _dense_batch(rng, it):
    if it == 1:  # new points farther from the centres than the tied points
        cc = (np.floor(rng.uniform(4, 8, (400, 3))) * 0.5 + 0.25).astype(np.float32)
        return cc + np.float32(0.24)
    return rng.uniform(-0.5, 4.5, (2000, 3)).astype(np.float32)


def test_delete_points(ikd):
    rng = np.random.default_rng(70)
    pts = uniform_cloud(rng, 20000)
    g, o = _pair(ikd, pts)
    sel = rng.choice(20000, 6000, replace=False)
    g.Delete_Points(pts[sel])
    o.Delete_Points(pts[sel])
    keep = np.ones(20000, bool)
    keep[sel] = False
    assert g.validnum() == o.validnum() == 14000
    assert_same_multiset(g.flatten()[0], pts[keep], "live set after Delete_Points")
    # duplicates inside one batch take the ordered path
    dup = np.concatenate([pts[keep][:10], pts[keep][:10]])
    g.Delete_Points(dup)
    o.Delete_Points(dup)
    assert g.validnum() == o.validnum() == 13990


def test_mixed_sequence_vs_oracle(ikd):
    rng = np.random.default_rng(80)
    g, o = _pair(ikd, uniform_cloud(rng, 20000, -30, 30), params=(0.5, 0.6, 0.5))
    next_id = 20000
    for it in range(8):
        b = uniform_cloud(rng, 2000, -32, 32)
        bid = np.arange(next_id, next_id + 2000, dtype=np.uint32)
        next_id += 2000
        assert g.Add_Points(b, it % 2 == 0, bid) == o.Add_Points(b, it % 2 == 0, bid)
        boxes = random_boxes(rng, 3, -30, 30, 2, 15)
        assert g.Delete_Point_Boxes(boxes) == o.Delete_Point_Boxes(boxes)
        live = o.flatten()[0]
        if it % 3 == 2:
            dp = live[rng.choice(live.shape[0], 500, replace=False)]
            g.Delete_Points(dp)
            o.Delete_Points(dp)
        assert g.validnum() == o.validnum()
        gl, gid = g.flatten()
        ol, oid = o.flatten()
        assert_same_multiset(gl, ol, f"live set step {it}")
        q = uniform_cloud(rng, 500, -32, 32)
        _, gd, gi, gc = g.knn(q, 5, 5.0)
        od = o.knn(q, 5, 5.0)
        assert knn_equivalent(gd, gi, gc, od[1], od[2], od[3]) == []
