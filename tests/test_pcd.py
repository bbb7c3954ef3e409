"""PCD files (include/ikd_pcd.hpp, ikd-tree_amd/pcd.py): the loader behind the
reference demos' pcl::io::loadPCDFile (examples/ikd_Tree_Search_demo.cpp:66).

The reference's demo file is not available, so the files here are synthetic,
written by an independent writer in each language and read by the other:
  - LZF decoding against streams worked out by hand from the format (literal
    runs, short and long back references, a reference that overlaps its own
    output) and malformed streams that must be rejected;
  - the Python writer's ascii / binary / binary_compressed files, with field
    types F4/F8/U1/U2/I2/I4, a COUNT > 1 field and a padding field, read by
    the C++ reader into PCL-layout point types (every value as float);
  - the C++ writer's three forms (its LZF encoder in the compressed one) read
    back by the Python reader, bit for bit.
"""
import importlib
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "include")
pcd = importlib.import_module("ikd-tree_amd.pcd")


def test_lzf_hand_streams():
    # literal run: c = n - 1 then n bytes
    assert pcd.lzf_decompress(bytes([2]) + b"abc", 3) == b"abc"
    # "abcabcabc": literal "abc", then length 6 (c >> 5 = 4) at distance 3 (low byte 2)
    assert pcd.lzf_decompress(bytes([2]) + b"abc" + bytes([4 << 5, 2]), 9) == b"abcabcabc"
    # ten 'a': literal 'a', then a reference of 9 at distance 1 overlapping its output
    assert pcd.lzf_decompress(bytes([0]) + b"a" + bytes([7 << 5, 0, 0]), 10) == b"a" * 10
    # long length: 7 + 20 + 2 = 29 copies of "xy" pattern at distance 2
    s = pcd.lzf_decompress(bytes([1]) + b"xy" + bytes([7 << 5, 20, 1]), 31)
    assert s == (b"xy" * 16)[:31]
    # distance with the high bits: 300 literal bytes, then 4 bytes from 257 back
    lit = bytes(range(256)) + bytes(range(44))
    stream = b"".join(bytes([31]) + lit[i:i + 32] for i in range(0, 288, 32)) + bytes([11]) + lit[288:]
    ref = bytes([(2 << 5) | 1, 0])  # length 4, distance ((1 << 8) | 0) + 1 = 257
    out = pcd.lzf_decompress(stream + ref, 304)
    assert out[:300] == lit and out[300:] == lit[300 - 257:300 - 257 + 4]
    for bad, n in ((bytes([5]) + b"ab", 6), (bytes([0]) + b"a" + bytes([1 << 5, 5]), 4), (bytes([2]) + b"abc", 4)):
        with pytest.raises(ValueError):
            pcd.lzf_decompress(bad, n)


def _mixed_fields(rng, n):
    return {
        "x": rng.uniform(-50, 50, n).astype(np.float32),
        "y": rng.uniform(-50, 50, n).astype(np.float64),
        "_": np.zeros(n, np.float32),
        "z": rng.integers(-30000, 30000, n).astype(np.int16),
        "rgb": rng.integers(0, 2 ** 31, n).astype(np.uint32),
        "intensity": rng.integers(0, 255, n).astype(np.uint8),
        "normal_x": rng.uniform(-1, 1, n).astype(np.float32),
        "normal_y": rng.integers(0, 60000, n).astype(np.uint16),
        "normal_z": rng.integers(-2 ** 30, 2 ** 30, n).astype(np.int32),
        "curvature": rng.uniform(0, 1, n).astype(np.float32),
        "hist": rng.uniform(0, 1, (n, 3)).astype(np.float32),
    }


READER = r"""
#include <cstdio>
#include "ikd_pcd.hpp"
#include "ikd_pcl_types.hpp"
int main(int argc, char** argv) {
    std::vector<ikd_pcl::PointXYZINormal> a;
    std::string err;
    if (ikd_pcd::load(argv[1], a, &err) != 0) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::FILE* f = std::fopen(argv[2], "wb");
    for (auto& p : a) {
        float r[8] = {p.x, p.y, p.z, p.intensity, p.normal_x, p.normal_y, p.normal_z, p.curvature};
        std::fwrite(r, 4, 8, f);
    }
    std::fclose(f);
    // the same file into the plain XYZ type
    std::vector<ikd_pcl::PointXYZ> b;
    if (ikd_pcd::loadPCDFile(argv[1], b) != 0 || b.size() != a.size()) return 2;
    for (size_t i = 0; i < a.size(); ++i) if (b[i].x != a[i].x || b[i].z != a[i].z) return 3;
    if (argc > 3) {  // write the three forms back with the C++ writer
        for (const char* m : {"ascii", "binary", "binary_compressed"})
            if (ikd_pcd::save(std::string(argv[3]) + "." + m + ".pcd", a, m) != 0) return 4;
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def reader(tmp_path_factory):
    d = tmp_path_factory.mktemp("pcd")
    src = d / "reader.cpp"
    src.write_text(READER)
    exe = d / "reader"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", INC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("mode", ["ascii", "binary", "binary_compressed"])
def test_cpp_reads_python_files(reader, tmp_path, mode):
    rng = np.random.default_rng(7)
    n = 700
    f = _mixed_fields(rng, n)
    path = tmp_path / f"m.{mode}.pcd"
    pcd.write_pcd(str(path), f, mode)
    out = tmp_path / "out.f32"
    r = subprocess.run([str(reader), str(path), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(n, 8)
    want = np.stack([f[k].astype(np.float64).astype(np.float32) for k in
                     ("x", "y", "z", "intensity", "normal_x", "normal_y", "normal_z", "curvature")], axis=1)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # and the Python reader agrees with what it wrote
    h, back = pcd.read_pcd(str(path))
    assert h["points"] == n
    for k, v in f.items():
        assert np.array_equal(back[k], v), k


def test_python_reads_cpp_files(reader, tmp_path):
    rng = np.random.default_rng(8)
    n = 2000
    f = {k: rng.uniform(-9, 9, n).astype(np.float32) for k in
         ("x", "y", "z", "intensity", "normal_x", "normal_y", "normal_z", "curvature")}
    # repeats give the C++ encoder back references to write (short and long)
    f["intensity"][:] = 3.5
    f["curvature"] = np.tile(f["curvature"][:50], n // 50)
    src = tmp_path / "src.pcd"
    pcd.write_pcd(str(src), f, "binary")
    r = subprocess.run([str(reader), str(src), str(tmp_path / "o.f32"), str(tmp_path / "cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for mode in ("ascii", "binary", "binary_compressed"):
        _, back = pcd.read_pcd(str(tmp_path / f"cpp.{mode}.pcd"))
        assert list(back) == ["x", "y", "z", "intensity", "normal_x", "normal_y", "normal_z", "curvature"]
        for k, v in f.items():
            assert np.array_equal(back[k].view(np.uint32), v.view(np.uint32)), (mode, k)
    comp = (tmp_path / "cpp.binary_compressed.pcd").stat().st_size
    assert comp < (tmp_path / "cpp.binary.pcd").stat().st_size  # the encoder used its back references


def test_load_xyz_and_bad_headers(tmp_path):
    p = tmp_path / "a.pcd"
    xyz = np.array([[1, 2, 3], [np.nan, 0, 0], [4, 5, 6]], np.float32)
    pcd.write_pcd(str(p), {"x": xyz[:, 0], "y": xyz[:, 1], "z": xyz[:, 2]}, "binary")
    assert pcd.load_xyz(str(p)).shape == (3, 3)
    assert np.array_equal(pcd.load_xyz(str(p), drop_nonfinite=True), xyz[[0, 2]])
    bad = tmp_path / "b.pcd"
    bad.write_bytes(b"VERSION 0.7\nFIELDS x y z\nSIZE 4 4\nTYPE F F F\nWIDTH 1\nPOINTS 1\nDATA ascii\n1 2 3\n")
    with pytest.raises(ValueError):
        pcd.read_pcd(str(bad))
