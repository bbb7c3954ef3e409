"""CPU checks of the drop-in boundary: the HIP library is built for gfx950,
loads, and exports every entry point include/ikd_tree_amd.h declares; the
Python mirror binds exactly that list.  No GPU call is made here."""
import ctypes
import importlib
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ikd_tree_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(ikd_\w+)\s*\(", text, flags=re.M)))


def _pkg():
    return importlib.import_module("ikd-tree_amd")


def test_header_declares_the_reference_api():
    names = _declared()
    for must in ("ikd_create", "ikd_build", "ikd_nearest_search", "ikd_box_search", "ikd_radius_search",
                 "ikd_add_points", "ikd_delete_points", "ikd_delete_point_boxes", "ikd_size", "ikd_validnum",
                 "ikd_tree_range", "ikd_root_alpha", "ikd_flatten"):
        assert must in names


def test_library_exports_every_declared_symbol():
    pkg = _pkg()
    assert os.path.exists(pkg.LIB_PATH), "build libikd_hip.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(pkg.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert sorted(pkg.EXPORTS) == _declared()


def test_library_carries_gfx950_code():
    # the embedded offload bundle names its target (this library is built here)
    data = open(_pkg().LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_product_path_does_not_import_the_oracle():
    pkg_dir = os.path.join(ROOT, "ikd-tree_amd")
    for dirpath, _, files in os.walk(pkg_dir):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "ikd_oracle" not in text, f


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    pkg = _pkg()
    monkeypatch.setattr(pkg, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        pkg.load_library()
    with pytest.raises(ImportError):
        pkg.KD_TREE()


def _compile_facade(out_path):
    src = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
    libdir = os.path.join(ROOT, "ikd-tree_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), src, "-L", libdir, "-likd_hip",
           f"-Wl,-rpath,{libdir}", "-o", str(out_path)]
    import subprocess
    return subprocess.run(cmd, capture_output=True, text=True)


def test_cpp_facade_compiles_and_links(tmp_path):
    """The header-only KD_TREE<PointType> facade builds with plain g++ against
    the C ABI, exactly as a reference caller would include ikd_Tree.h."""
    r = _compile_facade(tmp_path / "facade_check")
    assert r.returncode == 0, r.stderr


def test_served_result_buffers_are_checked():
    """Caller-owned result buffers must hold the request (rows, k, types); a
    mismatch is an error before anything reaches the library."""
    import numpy as np
    serving = importlib.import_module("ikd-tree_amd.serving")
    pts = np.zeros((10, 3), np.float32)
    outs = serving.result_buffers(16, 5)
    t = serving.Ticket(0, pts, 5, 5.0, 1, None, 0, outs)
    assert t.out_xyz.shape == (10, 5, 3) and t.out_counts.shape == (10,)
    assert t.req.out_d2 == outs[1].ctypes.data
    for bad in (serving.result_buffers(9, 5), serving.result_buffers(16, 4),
                (outs[0], outs[1].astype(np.float64), outs[2], outs[3])):
        with pytest.raises(ValueError):
            serving.Ticket(0, pts, 5, 5.0, 1, None, 0, bad)
