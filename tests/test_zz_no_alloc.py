"""GPU: the allocation-free batch kNN entry point
(ikd_batch_nearest_search_noalloc) against ikd_batch_nearest_search_device,
through the C ABI: identical counts, squared distances (bit for bit), ids and
points on two calls in a row, the second with fewer trees and queries (the
shared scratch keeps the first call's contents), plus its size limits."""
import ctypes
import importlib
import os

import numpy as np
import pytest

from common import INF, uniform_cloud

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DEV = os.environ.get("IKD_TEST_DEVICE", "cuda:0")


def _outs(n, kmax, dev):
    return (torch.empty((n, kmax, 3), dtype=torch.float32, device=dev),
            torch.empty((n, kmax), dtype=torch.float32, device=dev),
            torch.empty((n, kmax), dtype=torch.int32, device=dev),
            torch.empty((n,), dtype=torch.int32, device=dev))


def _inputs(rng, n_trees, nq):
    qt = rng.integers(0, n_trees, nq).astype(np.int32)
    q = uniform_cloud(rng, nq, -32, 32)
    k = rng.choice(np.array([1, 5, 10, 20], np.int32), nq)
    md = rng.choice(np.array([1.0, 5.0, INF]), nq)
    return qt, q, k, md


def test_noalloc_matches_the_allocating_entry_point():
    ikd = importlib.import_module("ikd-tree_amd")
    dev = torch.device(DEV)
    rng = np.random.default_rng(2025)
    trees = []
    for n in (20000, 3000, 70000, 500, 9000):
        t = ikd.KD_TREE(0.5, 0.6, 0.2)
        t.Build(uniform_cloud(rng, n, -30, 30))
        trees.append(t)
    b = ikd.Batch()
    kmax = 20
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    for call, (ntr, nq) in enumerate(((5, 40000), (2, 3000))):
        qt, q, k, md = _inputs(rng, ntr, nq)
        ref, got = _outs(nq, kmax, dev), _outs(nq, kmax, dev)
        for o in got:
            o.fill_(-7)
        b.nearest_search_device(trees[:ntr], d(qt), d(q), kmax, *ref, k=d(k), max_dist=d(md))
        b.synchronize()
        b.nearest_search_noalloc(trees[:ntr], d(qt), d(q), kmax, *got, k=d(k), max_dist=d(md))
        b.synchronize()
        rc, gc = ref[3].cpu().numpy(), got[3].cpu().numpy()
        assert np.array_equal(rc, gc), f"call {call}: counts"
        for i in range(nq):
            c = rc[i]
            assert np.array_equal(ref[1][i, :c].cpu().numpy().view(np.uint32),
                                  got[1][i, :c].cpu().numpy().view(np.uint32)), f"call {call} query {i}: d2"
        assert torch.equal(ref[2], got[2]) or all(
            np.array_equal(ref[2][i, :rc[i]].cpu().numpy(), got[2][i, :rc[i]].cpu().numpy()) for i in range(nq)), \
            f"call {call}: ids"
        assert torch.equal(ref[1], got[1]) and torch.equal(ref[2], got[2]) and torch.equal(ref[0], got[0]), \
            f"call {call}: full output rows"
    # one max_dist for every query (no per-query thresholds)
    qt, q, k, _ = _inputs(rng, 3, 5000)
    ref, got = _outs(5000, 10, dev), _outs(5000, 10, dev)
    b.nearest_search_device(trees[:3], d(qt), d(q), 10, *ref, max_dist=5.0)
    b.nearest_search_noalloc(trees[:3], d(qt), d(q), 10, *got, max_dist=5.0)
    b.synchronize()
    assert all(torch.equal(x, y) for x, y in zip(ref, got))


def test_noalloc_rejects_sizes_beyond_its_scratch():
    ikd = importlib.import_module("ikd-tree_amd")
    dev = torch.device(DEV)
    t = ikd.KD_TREE()
    t.Build(uniform_cloud(np.random.default_rng(1), 100))
    b = ikd.Batch()
    lib = ikd.load_library() if not hasattr(b, "_lib") else b._lib
    many = (ctypes.c_void_p * (ikd.IKD_NOALLOC_MAX_TREES + 1))(*([t._h.value] * (ikd.IKD_NOALLOC_MAX_TREES + 1)))
    qt = torch.zeros(4, dtype=torch.int32, device=dev)
    q = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    o = _outs(4, 5, dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    rc = lib.ikd_batch_nearest_search_noalloc(b._h, many, ikd.IKD_NOALLOC_MAX_TREES + 1, p(qt), p(q), 4, None, 5,
                                              None, 5.0, p(o[0]), p(o[1]), p(o[2]), p(o[3]))
    assert rc == ikd.IKD_ERR_CAPACITY
    assert b"IKD_NOALLOC_MAX_TREES" in lib.ikd_batch_last_error(b._h)
    one = (ctypes.c_void_p * 1)(t._h.value)
    nbig = ikd.IKD_NOALLOC_MAX_QUERIES + 1
    md = torch.zeros(1, dtype=torch.float64, device=dev)  # never read: the size check comes first
    rc = lib.ikd_batch_nearest_search_noalloc(b._h, one, 1, p(qt), p(q), nbig, None, 5, p(md), 5.0, p(o[0]),
                                              p(o[1]), p(o[2]), p(o[3]))
    assert rc == ikd.IKD_ERR_CAPACITY
