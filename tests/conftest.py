"""pytest configuration: registers the ``gpu`` marker and puts the repo root on sys.path."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def pytest_sessionstart(session):
    """IKD_TEST_LIB=<path> points the package at another build of the same
    sources (the CPU emulation build in tests/cpu_emu, for ASan runs of the
    gpu-marked tests on a machine without a GPU).

    On a GPU machine torch's HIP runtime is brought up here, before any test
    loads the tree library: torch ships its own HIP runtime, and it does not
    find the GPU once the library's runtime has opened the device first in the
    same process."""
    if os.environ.get("IKD_TEST_DEVICE", "cuda:0") != "cpu":
        try:
            import torch
            if torch.cuda.device_count() > 0:
                torch.zeros(1, device="cuda:0")
        except Exception:  # no torch, or no usable GPU: tests that need it skip or fail on their own
            pass
    path = os.environ.get("IKD_TEST_LIB")
    if path:
        import importlib
        pkg = importlib.import_module("ikd-tree_amd")
        pkg.LIB_PATH = os.path.abspath(path)
