"""Shared test helpers: seeded inputs, numpy brute force and result comparators.

The brute-force routines are independent of both the oracle and the HIP
library: they work on a flat array of live points, with the reference's
float32 evaluation order for squared distances (ikd-Tree/ikd_Tree.cpp:1373-1378)
and its half-open box rule (cpp:653-656, 1019-1026).
"""
from __future__ import annotations

import numpy as np

INF = float("inf")


# ----------------------------------------------------------------- inputs
def uniform_cloud(rng: np.random.Generator, n: int, lo: float = -50.0, hi: float = 50.0) -> np.ndarray:
    return (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)


def grid_cloud(rng: np.random.Generator, n: int, step: float = 0.25, extent: int = 80) -> np.ndarray:
    """Points on a coarse grid: many exact coordinate and distance ties."""
    return (rng.integers(-extent, extent, size=(n, 3)) * step).astype(np.float32)


def random_boxes(rng: np.random.Generator, nb: int, lo=-50.0, hi=50.0, smin=5.0, smax=20.0) -> np.ndarray:
    c = lo + (hi - lo) * rng.random((nb, 3))
    s = smin + (smax - smin) * rng.random((nb, 3))
    return np.concatenate([c - s / 2, c + s / 2], axis=1).astype(np.float32)


def grid_boxes(rng: np.random.Generator, nb: int, step: float = 0.25, extent: int = 80, smax: int = 40) -> np.ndarray:
    """Boxes whose faces lie on the grid of grid_cloud: points hit every face."""
    lo = rng.integers(-extent, extent, size=(nb, 3))
    sz = rng.integers(1, smax, size=(nb, 3))
    return (np.concatenate([lo, lo + sz], axis=1) * step).astype(np.float32)


# ------------------------------------------------------------ brute force
def d2_all(points: np.ndarray, q: np.ndarray) -> np.ndarray:
    """calc_dist for every point: ((dx*dx + dy*dy) + dz*dz), each op rounded to float32."""
    p = points.astype(np.float32, copy=False)
    dx = p[:, 0] - np.float32(q[0])
    dy = p[:, 1] - np.float32(q[1])
    dz = p[:, 2] - np.float32(q[2])
    return (dx * dx + dy * dy) + dz * dz


def max_dist_threshold(max_dist: float) -> float:
    """The float T with (double)d <= max_dist*max_dist  <=>  d <= T for every float d."""
    m = float(max_dist) * float(max_dist)
    t = np.float32(m)
    if float(t) > m:
        t = np.nextafter(t, np.float32(-np.inf))
    return float(t)


def brute_knn(points: np.ndarray, ids: np.ndarray, queries: np.ndarray, k: int, max_dist: float = INF,
              tie: str = "id"):
    """k smallest points per query among those with d2 <= max_dist^2 (double
    compare), ranked by (d2, id) or, with tie="xyz", by (d2, x, y, z, id)."""
    nq = queries.shape[0]
    m = float(max_dist) * float(max_dist)
    out_d = np.full((nq, k), np.inf, np.float32)
    out_id = np.full((nq, k), 0xFFFFFFFF, np.uint32)
    cnt = np.zeros(nq, np.int32)
    for i in range(nq):
        d = d2_all(points, queries[i])
        ok = np.nonzero(d.astype(np.float64) <= m)[0]
        if ok.size == 0:
            continue
        if tie == "xyz":
            p = points[ok]
            order = np.lexsort((ids[ok], p[:, 2], p[:, 1], p[:, 0], d[ok]))[:k]
        else:
            order = np.lexsort((ids[ok], d[ok]))[:k]
        sel = ok[order]
        c = sel.size
        cnt[i] = c
        out_d[i, :c] = d[sel]
        out_id[i, :c] = ids[sel]
    return out_d, out_id, cnt


def in_box(points: np.ndarray, box: np.ndarray) -> np.ndarray:
    """Half-open box membership: min <= p < max on every axis."""
    b = np.asarray(box, np.float32)
    return np.all((points >= b[:3]) & (points < b[3:]), axis=1)


def brute_radius(points: np.ndarray, q: np.ndarray, r: float):
    """Indices inside the ball (d2 <= r*r in float32) and a boundary band that the
    reference's sqrt-based subtree rule (cpp:1046-1087) may resolve either way."""
    d = d2_all(points, q)
    r32 = np.float32(r)
    inside = d <= r32 * r32
    dist = np.sqrt(d)
    band = np.abs(dist - r32) <= np.float32(4.0) * np.spacing(np.maximum(r32, dist)) + np.float32(4e-6) * r32
    return inside, band


# ------------------------------------------------------------- comparators
def sorted_rows(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
    if a.shape[0] == 0:
        return a
    order = np.lexsort((a[:, 2], a[:, 1], a[:, 0]))
    return a[order]


def assert_same_multiset(a: np.ndarray, b: np.ndarray, what: str = "point sets") -> None:
    sa, sb = sorted_rows(a), sorted_rows(b)
    assert sa.shape == sb.shape, f"{what}: sizes differ {sa.shape[0]} vs {sb.shape[0]}"
    if sa.size:
        eq = np.all(sa.view(np.uint32) == sb.view(np.uint32), axis=1)
        assert eq.all(), f"{what}: first mismatch at sorted row {int(np.argmin(eq))}: {sa[~eq][:3]} vs {sb[~eq][:3]}"


TIE_WINDOW = 1e-10  # the reference comparator's near-tie window, ikd_Tree.h:101


def knn_equivalent(d_a, id_a, c_a, d_b, id_b, c_b, strict_ties: bool = False):
    """Compare two batched kNN results query by query.

    Squared distances must agree bit-exactly as sorted lists, and the id sets
    must agree, except that entries tied (within the reference's 1e-10
    comparator window) with the k-th distance may be exchanged for other
    entries at that distance.  Returns the list of offending query indices."""
    bad = []
    c_a = np.asarray(c_a)
    c_b = np.asarray(c_b)
    for i in np.nonzero((c_a != c_b))[0]:
        bad.append(int(i))
    same_c = np.nonzero(c_a == c_b)[0]
    for i in same_c:
        c = int(c_a[i])
        if c == 0:
            continue
        da = np.sort(np.asarray(d_a[i, :c], np.float32))
        db = np.sort(np.asarray(d_b[i, :c], np.float32))
        sa = set(np.asarray(id_a[i, :c]).tolist())
        sb = set(np.asarray(id_b[i, :c]).tolist())
        if np.array_equal(da.view(np.uint32), db.view(np.uint32)) and sa == sb:
            continue
        if strict_ties:
            bad.append(int(i))
            continue
        kth = float(max(da[-1], db[-1]))
        inner_a = da[da.astype(np.float64) < kth - TIE_WINDOW]
        inner_b = db[db.astype(np.float64) < kth - TIE_WINDOW]
        if not np.array_equal(inner_a.view(np.uint32), inner_b.view(np.uint32)):
            bad.append(int(i))
            continue
        # ids of strictly-inner entries must agree
        ia = {int(x) for x, d in zip(id_a[i, :c], d_a[i, :c]) if float(d) < kth - TIE_WINDOW}
        ib = {int(x) for x, d in zip(id_b[i, :c], d_b[i, :c]) if float(d) < kth - TIE_WINDOW}
        if ia != ib:
            bad.append(int(i))
    return bad


def check_dump_invariants(d: np.ndarray, tight: bool = True) -> None:
    """Structural invariants of a pre-order node dump (oracle or device layout):
    TreeSize and invalid_point_num are the sums over the subtree (Update,
    ikd-Tree/ikd_Tree.cpp:1183-1200), tree_deleted holds exactly when no live
    point is left below, and the stored AABB bounds the live points (tight:
    equals their bounding box; otherwise contains it, since Delete_by_point
    leaves the deleted node's range as it was, cpp:717-721)."""
    import sys
    sys.setrecursionlimit(max(10000, sys.getrecursionlimit()))

    def rec(i):
        r = d[i]
        size, inv = 1, int(r["flags"] & 1)
        pts_live = [] if (r["flags"] & 1) else [(r["x"], r["y"], r["z"])]
        for c in (int(r["left"]), int(r["right"])):
            if c >= 0:
                s, iv, pl = rec(c)
                size += s
                inv += iv
                pts_live += pl
        assert size == r["tree_size"], f"row {i}: TreeSize"
        assert inv == r["invalid"], f"row {i}: invalid_point_num"
        td = bool(r["flags"] & 2)
        assert td == (len(pts_live) == 0), f"row {i}: tree_deleted <=> no live point"
        if pts_live and tight is not None:
            a = np.array(pts_live, np.float32)
            bb = np.concatenate([a.min(axis=0), a.max(axis=0)])
            if tight:
                assert np.array_equal(bb, r["range"]), f"row {i}: AABB is the tight live bbox"
            else:
                assert np.all(r["range"][:3] <= bb[:3]) and np.all(r["range"][3:] >= bb[3:]), f"row {i}: AABB"
        return size, inv, pts_live

    if d.shape[0]:
        rec(0)
