"""GPU: the C++ drop-in facade (include/ikd_Tree_amd.hpp) driven like the
reference's demo, checked against the oracle on the same inputs.  The point
type carries an intensity payload that must round-trip."""
import os
import subprocess

import numpy as np
import pytest

from common import assert_same_multiset, in_box, random_boxes, uniform_cloud
from oracle import OracleTree
from test_abi import _compile_facade

pytestmark = pytest.mark.gpu


def test_cpp_facade_matches_oracle(tmp_path):
    exe = tmp_path / "facade_check"
    r = _compile_facade(exe)
    assert r.returncode == 0, r.stderr
    rng = np.random.default_rng(555)
    pts = uniform_cloud(rng, 20000, -20, 20)
    qs = uniform_cloud(rng, 400, -21, 21)
    boxes = random_boxes(rng, 4, -20, 20, 3, 10)
    pts.tofile(tmp_path / "p.f32")
    qs.tofile(tmp_path / "q.f32")
    boxes.tofile(tmp_path / "b.f32")
    out = subprocess.run([str(exe), str(tmp_path / "p.f32"), str(tmp_path / "q.f32"), str(tmp_path / "b.f32"),
                          str(tmp_path / "o")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    o = OracleTree(0.5, 0.6, 0.5)
    o.Build(pts)
    # per-point Nearest_Search
    knn = np.fromfile(tmp_path / "o_knn.f32", np.float32).reshape(400, 5, 5)
    _, od, oid, oc = o.knn(qs, 5)
    for i in range(400):
        c = int(oc[i])
        assert np.array_equal(knn[i, :c, 4], od[i, :c])
        got = knn[i, :c, 3].astype(np.int64)  # intensity = input index of the returned point
        assert np.array_equal(np.sort(got), np.sort(oid[i, :c].astype(np.int64)))
        assert np.array_equal(knn[i, :c, :3], pts[got])
    box = np.fromfile(tmp_path / "o_box.f32", np.float32).reshape(-1, 4)
    assert_same_multiset(box[:, :3], pts[in_box(pts, boxes[0])], "facade Box_Search")
    assert np.array_equal(box[:, :3], pts[box[:, 3].astype(np.int64)])
    rad = np.fromfile(tmp_path / "o_radius.f32", np.float32).reshape(-1, 4)
    assert_same_multiset(rad[:, :3], o.Radius_Search(qs[0], 7.5)[0], "facade Radius_Search")
    n_del = o.Delete_Point_Boxes(boxes)
    n_add = o.Add_Points(qs, True)
    line = out.stdout.strip().split()
    assert int(line[1]) == n_del and int(line[3]) == n_add and int(line[5]) == o.validnum()
    flat = np.fromfile(tmp_path / "o_flat.f32", np.float32).reshape(-1, 4)
    assert_same_multiset(flat[:, :3], o.flatten()[0], "facade flatten after updates")


def _build_exe(src, exe):
    libdir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ikd-tree_amd")
    inc = os.path.join(os.path.dirname(libdir), "include")
    return subprocess.run(["g++", "-std=c++17", "-O2", "-I", inc, os.path.join(os.path.dirname(__file__), "cpp", src),
                           "-L", libdir, "-likd_hip", f"-Wl,-rpath,{libdir}", "-lpthread", "-o", str(exe)],
                          capture_output=True, text=True)


def test_concurrent_nearest_search_on_one_tree(tmp_path):
    """8 host threads call Nearest_Search on one tree at once (FAST-LIO's
    OpenMP loop); every result equals the single-threaded one."""
    exe = tmp_path / "concurrent_search"
    r = _build_exe("concurrent_search.cpp", exe)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), "200000", "4000", "8"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")


def test_reference_demo_call_sequences(tmp_path):
    """The three reference demos' call sequences, typed against the facade,
    give the oracle's results on the GPU."""
    from demo_harness import check_outputs, write_inputs
    exe = tmp_path / "demo_calls"
    r = _build_exe("demo_calls.cpp", exe)
    assert r.returncode == 0, r.stderr
    ind, outd = tmp_path / "in", tmp_path / "out"
    ind.mkdir()
    outd.mkdir()
    inputs = write_inputs(ind, n=20000)
    out = subprocess.run([str(exe), str(ind), str(outd)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    check_outputs(inputs, outd, out.stdout)


def test_pcd_map_through_the_facade(tmp_path):
    """The Search demo's flow without PCL on the GPU: a 200k-point PCD map
    loaded into PointXYZINormal records, built, searched; results keep their
    full records, distances equal brute force, the box is exact."""
    from test_emu_host import _pcd_map
    _pcd_map(tmp_path / "map.pcd", 200000, 32)
    exe = tmp_path / "pcd_payload"
    r = _build_exe("pcd_payload.cpp", exe)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), str(tmp_path / "map.pcd")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
