"""GPU: the request loop over many device-resident maps (ikd_server_*).

Bar: every request's kNN results (counts, squared distances bit for bit, ids,
points), Add_Points return value and trim count, and every map's final live
point set, are identical when the requests run mixed and overlapped in the
server and when the same requests run alone, one after another, through the
plain entry points.
"""
import importlib

import numpy as np
import pytest

from common import assert_same_multiset

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ikd():
    return importlib.import_module("ikd-tree_amd")


@pytest.fixture(scope="module")
def serving():
    return importlib.import_module("ikd-tree_amd.serving")


@pytest.fixture(scope="module")
def scenes():
    return importlib.import_module("ikd-tree_amd.scenes")


def _trim_box(scene, frame, half=60.0):
    """Six half-open slabs: everything outside a cube of side 2*half around the sensor."""
    p = scene.pose(frame)
    big = 1e7
    lo = p - half
    hi = p + half
    return np.array([
        [-big, -big, -big, lo[0], big, big], [hi[0], -big, -big, big, big, big],
        [-big, -big, -big, big, lo[1], big], [-big, hi[1], -big, big, big, big],
        [-big, -big, -big, big, big, lo[2]], [-big, -big, hi[2], big, big, big],
    ], np.float32)


@pytest.mark.parametrize("defer", [None, 5000])
def test_mixed_equals_alone(ikd, serving, scenes, defer):
    n_maps, per_map = 4, 5
    sc = [scenes.MappingScene(1005 + m) for m in range(n_maps)]
    seeds = [np.concatenate([s.scan(f) for f in range(2)]) for s in sc]
    reqs = []  # submission order, interleaved over the maps
    for j in range(per_map):
        for m in range(n_maps):
            pts = sc[m].scan(2 + j)[: 8000 + 3000 * m]  # uneven sizes
            trim = _trim_box(sc[m], 2 + j) if j == 3 else None
            reqs.append((m, pts, trim))

    trees = []
    for m in range(n_maps):
        t = ikd.KD_TREE(0.5, 0.6, 0.5)
        t.Build(seeds[m])
        trees.append(t)
    srv = serving.Server(slots=4, hbm_budget_bytes=8 << 30)
    ids = [srv.add_map(t, defer_rebuild_points=defer) for t in trees]
    # requests write into their own arrays, into caller-owned buffers larger
    # than they need, or into page-locked caller buffers (filled by DMA)
    tickets = [srv.submit(ids[m], pts, 5, 5.0, 1, trim, tag=i,
                          outs=None if i % 3 == 0 else serving.result_buffers(len(pts) + 97, 5, i % 3 == 2))
               for i, (m, pts, trim) in enumerate(reqs)]
    assert [t.req.flags for t in tickets[:3]] == [0, 0, serving.REQ_PINNED_RESULTS]
    done = srv.drain()
    srv.close()
    assert len(done) == len(tickets)
    assert all(t.status == 0 for t in tickets), [t.status for t in tickets]
    assert all(t.req.t_done_ns >= t.req.t_start_ns >= t.req.t_arrival_ns for t in tickets)

    alone = []
    for m in range(n_maps):
        t = ikd.KD_TREE(0.5, 0.6, 0.5)
        t.Build(seeds[m])
        alone.append(t)
    for i, (m, pts, trim) in enumerate(reqs):
        x, d, idx, c = alone[m].knn(pts, 5, 5.0)
        tk = tickets[i]
        assert np.array_equal(tk.out_counts, c), f"request {i}: counts"
        assert np.array_equal(tk.out_d2.view(np.uint32), d.view(np.uint32)), f"request {i}: d2"
        assert np.array_equal(tk.out_ids, idx), f"request {i}: ids"
        assert np.array_equal(tk.out_xyz, x), f"request {i}: points"
        added = alone[m].Add_Points(pts, True)
        assert tk.req.added == added, f"request {i}: Add_Points return"
        if trim is not None:
            assert tk.req.trimmed == alone[m].Delete_Point_Boxes(trim), f"request {i}: trim count"
    for m in range(n_maps):
        a, _ = trees[m].flatten()
        b, _ = alone[m].flatten()
        assert_same_multiset(a, b)


def test_bad_requests_rejected(ikd, serving):
    t = ikd.KD_TREE()
    t.Build(np.zeros((1, 3), np.float32))
    srv = serving.Server(slots=1, hbm_budget_bytes=1 << 30)
    m = srv.add_map(t)
    with pytest.raises(ikd.IkdError):
        srv.submit(m + 1, np.zeros((4, 3), np.float32))
    with pytest.raises(ikd.IkdError):
        srv.submit(m, np.zeros((4, 3), np.float32), k=65)
    srv.close()


def test_partial_and_empty_requests(ikd, serving, scenes):
    """Requests that skip a stage: an empty scan, an insert without kNN
    (k = 0), a kNN without insert (downsample -1), a trim alone; with own,
    caller and page-locked result buffers.  Each matches the same calls run
    alone through the plain entry points."""
    sc = scenes.MappingScene(1009)
    seed = np.concatenate([sc.scan(f) for f in range(2)])
    trees = []
    for _ in range(2):
        t = ikd.KD_TREE(0.5, 0.6, 0.5)
        t.Build(seed)
        trees.append(t)
    srv = serving.Server(slots=2, hbm_budget_bytes=4 << 30)
    m = srv.add_map(trees[0])
    s2, s3, s4 = sc.scan(2)[:9000], sc.scan(3)[:7000], sc.scan(4)[:5000]
    trim = _trim_box(sc, 4, half=40.0)
    reqs = [  # (points, k, downsample, trim, outs)
        (np.zeros((0, 3), np.float32), 5, 1, None, None),
        (s2, 0, 1, None, None),
        (s3, 5, -1, None, serving.result_buffers(7000, 5, True)),
        (s3, 5, -1, None, serving.result_buffers(7100, 5, False)),
        (np.zeros((0, 3), np.float32), 0, -1, trim, None),
        (s4, 5, 1, None, serving.result_buffers(5000, 5, True)),
    ]
    tickets = [srv.submit(m, p, k, 5.0, ds, tr, tag=i, outs=o) for i, (p, k, ds, tr, o) in enumerate(reqs)]
    srv.drain()
    srv.close()
    assert all(t.status == 0 for t in tickets), [t.status for t in tickets]
    alone = trees[1]
    for i, (p, k, ds, tr, o) in enumerate(reqs):
        tk = tickets[i]
        if k > 0 and p.shape[0] > 0:
            x, d, idx, c = alone.knn(p, k, 5.0)
            assert np.array_equal(tk.out_counts, c), f"request {i}: counts"
            assert np.array_equal(tk.out_d2.view(np.uint32), d.view(np.uint32)), f"request {i}: d2"
            assert np.array_equal(tk.out_ids, idx), f"request {i}: ids"
            assert np.array_equal(tk.out_xyz, x), f"request {i}: points"
        if ds >= 0 and p.shape[0] > 0:
            assert tk.req.added == alone.Add_Points(p, bool(ds)), f"request {i}: Add_Points return"
        else:
            assert tk.req.added == 0
        if tr is not None:
            assert tk.req.trimmed == alone.Delete_Point_Boxes(tr), f"request {i}: trim count"
    a, _ = trees[0].flatten()
    b, _ = alone.flatten()
    assert_same_multiset(a, b)


def test_budget_counts_resident_maps(ikd, serving, scenes):
    """The admission budget covers the maps' own device memory: with room for
    the resident maps plus one request, requests for four different maps run
    one at a time (each starts after the previous one finished); with a large
    budget the same requests finish too.  The resident figure follows the maps'
    growth and equals the sum of their device_bytes()."""
    n_maps = 4
    sc = [scenes.MappingScene(1105 + m) for m in range(n_maps)]
    scans = [[s.scan(f)[:6000] for f in range(4)] for s in sc]
    for budget_kind in ("tight", "large"):
        trees = []
        for m in range(n_maps):
            t = ikd.KD_TREE(0.5, 0.6, 0.5)
            t.Build(scans[m][0])
            trees.append(t)
        probe = serving.Server(slots=1, hbm_budget_bytes=1 << 40)
        for t in trees:
            probe.add_map(t)
        resident0, _ = probe.memory()
        probe.close()
        assert resident0 == sum(t.device_bytes() for t in trees)
        one = serving.Server.request_bytes(6000, 5, 1)
        budget = resident0 + one + (1 << 20) if budget_kind == "tight" else 64 << 30
        srv = serving.Server(slots=4, hbm_budget_bytes=budget)
        ids = [srv.add_map(t) for t in trees]
        tickets = [srv.submit(ids[m], scans[m][j], 5, 5.0, 1, tag=j) for j in range(1, 4) for m in range(n_maps)]
        srv.drain()
        srv.wait_idle()
        resident, running = srv.memory()
        srv.close()
        assert all(t.status == 0 for t in tickets)
        assert running == 0
        assert resident == sum(t.device_bytes() for t in trees) >= resident0
        if budget_kind == "tight":
            order = sorted(tickets, key=lambda t: t.req.t_start_ns)
            for a, b in zip(order, order[1:]):
                assert b.req.t_start_ns >= a.req.t_done_ns, "two requests ran at once under a one-request budget"
        for t in trees:
            t.close()
