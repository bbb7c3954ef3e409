"""Inputs, runner and oracle comparison for tests/cpp/demo_calls.cpp (the
reference demos' call sequences against the C++ facade).  Used by the CPU
emulation test (tests/test_emu_host.py) and the GPU test
(tests/test_gpu_facade.py) with the library each of them links."""
from __future__ import annotations

import os
import subprocess

import numpy as np

from common import assert_same_multiset, brute_knn, knn_equivalent, random_boxes, uniform_cloud
from oracle import OracleTree

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROUNDS = 4


def compile_demo(exe, libdir, libname, extra=()):
    cmd = ["g++", "-std=c++17", "-O1", "-g", *extra, "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "cpp", "demo_calls.cpp"), "-o", str(exe), "-L", libdir, f"-l{libname}",
           f"-Wl,-rpath,{libdir}"]
    return subprocess.run(cmd, capture_output=True, text=True)


def write_inputs(d, n=6000, seed=2024):
    rng = np.random.default_rng(seed)
    cloud = uniform_cloud(rng, n, -20, 20)
    cloud.tofile(d / "cloud.f32")
    np.array([ROUNDS], np.float32).tofile(d / "rounds.f32")
    inputs = {"cloud": cloud, "rounds": []}
    live = [cloud]
    for r in range(ROUNDS):
        inc = uniform_cloud(rng, 300, -20, 20)
        pool = np.concatenate(live + [inc])
        dec = pool[rng.choice(pool.shape[0], 200, replace=False)]
        delbox = random_boxes(rng, 4, -20, 20, 3, 8)
        addbox = random_boxes(rng, 4, -20, 20, 3, 8)
        target = uniform_cloud(rng, 100, -20, 20)
        for name, arr in (("inc", inc), ("dec", dec), ("delbox", delbox), ("addbox", addbox), ("target", target)):
            arr.tofile(d / f"{name}{r}.f32")
        inputs["rounds"].append(dict(inc=inc, dec=dec, delbox=delbox, addbox=addbox, target=target))
        live.append(inc)
    searchbox = np.array([[-5, -5, -5, 6, 7, 8]], np.float32)
    ball = np.array([1.0, -2.0, 3.0, 6.5], np.float32)
    asyncbox = np.array([[0, -10, -20, 10, 0, 20]], np.float32)
    searchbox.tofile(d / "searchbox.f32")
    ball.tofile(d / "ball.f32")
    asyncbox.tofile(d / "asyncbox.f32")
    inputs.update(searchbox=searchbox, ball=ball, asyncbox=asyncbox)
    return inputs


def check_outputs(inputs, outdir, stdout):
    """Replays the demos' sequences on the oracle and compares: validnum and
    box-delete counts every round, each Nearest_Search (squared distances and
    points; after an Add_Point_Boxes round a query may differ only where the
    oracle's own walk missed a restored point behind a stale range, and then
    the facade's answer must be the exact one), the live set at the end, the
    removed points (each one a point that was deleted and is no longer live),
    and the Search and Async demos' results."""
    o = OracleTree(0.3, 0.6, 0.2)
    o.Build(inputs["cloud"])
    log = np.fromfile(outdir / "demo_rounds.f32", np.float32).reshape(ROUNDS, 4)
    knn = np.fromfile(outdir / "demo_knn.f32", np.float32).reshape(ROUNDS, 100, 5, 4)
    removed = np.fromfile(outdir / "demo_removed.f32", np.float32).reshape(-1, 3)
    for r, rd in enumerate(inputs["rounds"]):
        o.Add_Points(rd["inc"], False)
        o.Delete_Points(rd["dec"])
        if r % 2 == 1:
            assert int(log[r, 2]) == o.Delete_Point_Boxes(rd["delbox"]), f"round {r}: Delete_Point_Boxes count"
        if r % 2 == 0:
            o.Add_Point_Boxes(rd["addbox"])
        assert int(log[r, 0]) == o.validnum(), f"round {r}: validnum"
        _, od, oi, oc = o.knn(rd["target"], 5)
        lx, li = o.flatten()
        gd = knn[r, :, :, 3].copy()
        gc = (gd >= 0).sum(axis=1).astype(np.int32)
        gd[gd < 0] = np.inf
        # identities by coordinates: map returned points to live ids
        lut = {tuple(p.tolist()): int(i) for p, i in zip(lx, li)}
        gi = np.full((100, 5), 0xFFFFFFFF, np.uint32)
        for q in range(100):
            for j in range(gc[q]):
                gi[q, j] = lut[tuple(knn[r, q, j, :3].tolist())]
        bad = knn_equivalent(gd, gi, gc, od, oi, oc)
        if bad:
            bd, bi, bc = brute_knn(lx, li, rd["target"], 5, tie="xyz")
            for q in bad:
                assert knn_equivalent(od[q:q + 1], oi[q:q + 1], oc[q:q + 1], bd[q:q + 1], bi[q:q + 1], bc[q:q + 1]), \
                    f"round {r} query {q}: facade differs from an exact oracle answer"
                assert not knn_equivalent(gd[q:q + 1], gi[q:q + 1], gc[q:q + 1], bd[q:q + 1], bi[q:q + 1],
                                          bc[q:q + 1]), f"round {r} query {q}: neither the oracle's nor exact"
    live = np.fromfile(outdir / "demo_live.f32", np.float32).reshape(-1, 3)
    assert_same_multiset(live, o.flatten()[0], "demo: live set at the end")
    # removed points were all deleted: none of them is live at the end
    if removed.size:
        live_set = {tuple(p.tolist()) for p in live}
        assert not any(tuple(p.tolist()) in live_set for p in removed)
    # Search demo
    s = OracleTree(0.3, 0.6, 0.2)
    s.Build(inputs["cloud"])
    sb = np.fromfile(outdir / "search_box.f32", np.float32).reshape(-1, 3)
    assert_same_multiset(sb, s.Box_Search(inputs["searchbox"][0])[0], "Search demo: Box_Search")
    sr = np.fromfile(outdir / "search_radius.f32", np.float32).reshape(-1, 3)
    ball = inputs["ball"]
    assert_same_multiset(sr, s.Radius_Search(ball[:3], float(ball[3]))[0], "Search demo: Radius_Search")
    # Async demo
    a = OracleTree(0.3, 0.6, 0.2)
    a.Build(inputs["cloud"])
    nd = a.Delete_Point_Boxes(inputs["asyncbox"])
    al = np.fromfile(outdir / "async_live.f32", np.float32).reshape(-1, 3)
    assert_same_multiset(al, a.flatten()[0], "Async demo: flatten(Root_Node, PCL_Storage, NOT_RECORD)")
    line = [ln for ln in stdout.splitlines() if ln.startswith("async:")][0].split()
    assert int(line[2]) == nd
