"""GPU: many trees' kNN queries in one launch (ikd_batch_nearest_search_device).

Bars: every query's result (counts, squared distances bit for bit, ids) is
identical to the same query run alone against its own tree through the plain
entry point, and equal to brute force under the (d2, x, y, z, id) order; uneven
tree sizes, uneven query counts, mixed k and mixed max_dist in one launch.
"""
import importlib
import os

import numpy as np
import pytest

from common import INF, brute_knn, grid_cloud, uniform_cloud

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
# the device the tensors live on ("cpu" for the CPU emulation build of the library)
DEV = os.environ.get("IKD_TEST_DEVICE", "cuda:0")


@pytest.fixture(scope="module")
def ikd():
    return importlib.import_module("ikd-tree_amd")


def _outs(n, kmax, dev):
    return (torch.empty((n, kmax, 3), dtype=torch.float32, device=dev),
            torch.empty((n, kmax), dtype=torch.float32, device=dev),
            torch.empty((n, kmax), dtype=torch.int32, device=dev),
            torch.empty((n,), dtype=torch.int32, device=dev))


def _mixed_batch(seed, sizes, nqs, ks=(1, 5, 10, 20), dists=(1.0, 5.0, INF), grid_tree=None):
    rng = np.random.default_rng(seed)
    clouds = []
    for i, n in enumerate(sizes):
        clouds.append(grid_cloud(rng, n) if i == grid_tree else uniform_cloud(rng, n, -30, 30))
    qt, qs, qk, qd = [], [], [], []
    for i, m in enumerate(nqs):
        qs.append(uniform_cloud(rng, m, -32, 32))
        qt.append(np.full(m, i, np.int32))
        qk.append(rng.choice(ks, m).astype(np.int32))
        qd.append(rng.choice(dists, m).astype(np.float64))
    perm = rng.permutation(sum(nqs))  # interleave the trees' queries
    cat = lambda xs: np.concatenate(xs)[perm]
    return clouds, cat(qt), cat(qs), cat(qk), cat(qd)


@pytest.mark.parametrize("lanes", [-1, 1, 2, 0])
def test_batch_matches_single_tree_calls(ikd, lanes):
    dev = torch.device(DEV)
    sizes = [1, 17, 900, 5000, 40000, 120000]
    nqs = [30, 200, 1000, 3000, 700, 9000]
    clouds, qt, q, qk, qd = _mixed_batch(11, sizes, nqs, grid_tree=3)
    trees = []
    for c in clouds:
        t = ikd.KD_TREE(0.5, 0.6, 0.2)
        t.Build(c)
        trees.append(t)
    kmax = 20
    b = ikd.Batch()
    b.set_tuning(ikd.IKD_TUNE_LANES_PER_QUERY, lanes)
    o = _outs(q.shape[0], kmax, dev)
    b.nearest_search_device(trees, torch.from_numpy(qt).to(dev), torch.from_numpy(q).to(dev), kmax, *o,
                            k=torch.from_numpy(qk).to(dev), max_dist=torch.from_numpy(qd).to(dev))
    b.synchronize()
    bx, bd, bi, bc = (x.cpu().numpy() for x in o)
    bi = bi.view(np.uint32)
    for ti, t in enumerate(trees):
        for k in (1, 5, 10, 20):
            for md in (1.0, 5.0, INF):
                sel = np.nonzero((qt == ti) & (qk == k) & (qd == md))[0]
                if sel.size == 0:
                    continue
                sx, sd, si, sc = t.knn(q[sel], k, md)
                assert np.array_equal(bc[sel], sc), f"counts tree {ti} k {k} md {md}"
                assert np.array_equal(bd[sel, :k].view(np.uint32), sd.view(np.uint32))
                assert np.array_equal(bi[sel, :k], si)
                assert np.array_equal(bx[sel, :k], sx)
                # slots past the query's k are empty
                assert np.all(np.isinf(bd[sel, k:])) and np.all(bi[sel, k:] == 0xFFFFFFFF)
                if ti in (0, 1, 2):  # brute force on the small trees
                    ids = np.arange(clouds[ti].shape[0], dtype=np.uint32)
                    rd, ri, rc = brute_knn(clouds[ti], ids, q[sel], k, md, tie="xyz")
                    assert np.array_equal(sc, rc)
                    assert np.array_equal(sd.view(np.uint32), rd.view(np.uint32))
                    assert np.array_equal(si, ri)
    for t in trees:
        t.close()


def test_batch_uniform_k_and_bad_tree_index(ikd):
    dev = torch.device(DEV)
    rng = np.random.default_rng(5)
    a, c = ikd.KD_TREE(), ikd.KD_TREE()
    pa, pc = uniform_cloud(rng, 3000), uniform_cloud(rng, 500)
    a.Build(pa)
    c.Build(pc)
    q = uniform_cloud(rng, 400)
    qt = rng.integers(-1, 3, 400).astype(np.int32)  # -1 and 2 name no tree
    o = _outs(400, 5, dev)
    b = ikd.Batch()
    b.nearest_search_device([a, c], torch.from_numpy(qt).to(dev), torch.from_numpy(q).to(dev), 5, *o,
                            max_dist=7.5)
    b.synchronize()
    bd, bi, bc = o[1].cpu().numpy(), o[2].cpu().numpy().view(np.uint32), o[3].cpu().numpy()
    bad = (qt < 0) | (qt > 1)
    assert np.all(bc[bad] == 0) and np.all(np.isinf(bd[bad])) and np.all(bi[bad] == 0xFFFFFFFF)
    for ti, t in enumerate((a, c)):
        sel = np.nonzero(qt == ti)[0]
        _, sd, si, sc = t.knn(q[sel], 5, 7.5)
        assert np.array_equal(bc[sel], sc)
        assert np.array_equal(bd[sel].view(np.uint32), sd.view(np.uint32))
        assert np.array_equal(bi[sel], si)


def test_batch_orders_after_tree_updates(ikd):
    """A launch sees every update queued on the trees before it, without a host sync."""
    dev = torch.device(DEV)
    rng = np.random.default_rng(9)
    t = ikd.KD_TREE(0.5, 0.6, 0.5)
    t.Build(uniform_cloud(rng, 20000))
    extra = torch.from_numpy(uniform_cloud(rng, 5000)).to(dev)
    t.add_points_device(extra, False)
    q = uniform_cloud(rng, 2000)
    o = _outs(2000, 5, dev)
    b = ikd.Batch()
    b.nearest_search_device([t], torch.zeros(2000, dtype=torch.int32, device=dev), torch.from_numpy(q).to(dev), 5,
                            *o)
    b.synchronize()
    _, sd, si, sc = t.knn(q, 5)
    assert np.array_equal(o[1].cpu().numpy().view(np.uint32), sd.view(np.uint32))
    assert np.array_equal(o[2].cpu().numpy().view(np.uint32), si)


def test_batch_argument_errors(ikd):
    dev = torch.device(DEV)
    t = ikd.KD_TREE()
    t.Build(uniform_cloud(np.random.default_rng(1), 100))
    b = ikd.Batch()
    q = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    qt = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(ikd.IkdError):
        b.nearest_search_device([t], qt, q, 0, *_outs(4, 1, dev))
    with pytest.raises(ikd.IkdError):
        b.nearest_search_device([t], qt, q, 65, *_outs(4, 65, dev))
    with pytest.raises(ikd.IkdError):
        b.nearest_search_device([], qt, q, 5, *_outs(4, 5, dev))


def test_batch_radius_matches_single_tree_calls(ikd):
    dev = torch.device(DEV)
    rng = np.random.default_rng(21)
    sizes = [50, 3000, 60000]
    trees, clouds = [], []
    for n in sizes:
        c = uniform_cloud(rng, n, -20, 20)
        t = ikd.KD_TREE()
        t.Build(c)
        trees.append(t)
        clouds.append(c)
    nq = 3000
    qt = rng.integers(-1, 3, nq).astype(np.int32)
    cen = uniform_cloud(rng, nq, -22, 22)
    rad = (0.2 + 2.8 * rng.random(nq)).astype(np.float32)
    b = ikd.Batch()
    off = torch.empty(nq + 1, dtype=torch.int64, device=dev)
    args = (trees, torch.from_numpy(qt).to(dev), torch.from_numpy(cen).to(dev), torch.from_numpy(rad).to(dev), off)
    total = b.radius_search_device(*args)
    xyz = torch.empty((max(total, 1), 3), dtype=torch.float32, device=dev)
    ids = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    assert b.radius_search_device(*args, xyz, ids) == total
    b.synchronize()
    o = off.cpu().numpy()
    gx, gi = xyz.cpu().numpy(), ids.cpu().numpy().view(np.uint32)
    bad = (qt < 0) | (qt > 2)
    assert np.all(o[1:][bad] == o[:-1][bad])
    for ti, t in enumerate(trees):
        sel = np.nonzero(qt == ti)[0]
        so, sx, si = t.radius_search_batch(cen[sel], rad[sel])
        for j, qi in enumerate(sel):
            a = np.sort(gi[o[qi]:o[qi + 1]])
            c = np.sort(si[so[j]:so[j + 1]])
            assert np.array_equal(a, c), f"tree {ti} query {qi}"
    # too small an output buffer: IKD_ERR_CAPACITY, counts still valid
    if total > 1:
        with pytest.raises(ikd.IkdError):
            b.radius_search_device(*args, xyz[: total - 1], ids[: total - 1])


@pytest.mark.parametrize("per_query", [False, True])
def test_recorded_launch_replays(ikd, per_query):
    """The multi-tree launch recorded into a HIP graph (Batch.knn_graph) gives
    the entry point's results at every replay: the first, one after new
    queries were written into the recorded buffer, and one after a tree grew
    (its node pool moved).  Eager work and host synchronizations run between
    the replays, and every replay is checked."""
    dev = torch.device(DEV)
    clouds, qt, q, qk, qd = _mixed_batch(21, [300, 8000, 60000], [500, 1500, 2500])
    trees = []
    for c in clouds:
        t = ikd.KD_TREE(0.5, 0.6, 0.5)
        t.Build(c)
        trees.append(t)
    kmax = 20
    n = q.shape[0]
    d_qt, d_q = torch.from_numpy(qt).to(dev), torch.from_numpy(q).to(dev)
    d_k = torch.from_numpy(qk).to(dev)
    d_md = torch.from_numpy(qd).to(dev) if per_query else 5.0
    b = ikd.Batch()
    rec = _outs(n, kmax, dev)
    g = b.knn_graph(trees, d_qt, d_q, kmax, *rec, k=d_k, max_dist=d_md)
    rng = np.random.default_rng(22)

    def check(step):
        ref = _outs(n, kmax, dev)
        b.nearest_search_device(trees, d_qt, d_q, kmax, *ref, k=d_k, max_dist=d_md)
        b.synchronize()
        for x in rec:
            x.fill_(-7)  # the replay must write every slot
        g.launch()
        b.synchronize()
        for a_, r_ in zip(rec, ref):
            assert torch.equal(a_.cpu(), r_.cpu()), f"replay {step}"

    check(0)
    d_q.copy_(torch.from_numpy(uniform_cloud(rng, n, -32, 32)).to(dev))
    check(1)
    big = uniform_cloud(rng, 150000, -30, 30)
    trees[0].Add_Points(big, False)  # the pool of the 300-point tree grows (and moves)
    check(2)
    check(3)
    g.close()
    for t in trees:
        t.close()
