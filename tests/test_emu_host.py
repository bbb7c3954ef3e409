"""CPU checks of the library's host code on the emulation build
(tests/cpu_emu: the same HIP sources compiled by g++ with AddressSanitizer,
kernels run on host threads).  No GPU is needed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")


@pytest.fixture(scope="module")
def emu_lib():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    subprocess.run(["make", "-s", "-C", EMU], check=True)
    return os.path.join(EMU, "libikd_emu.so")


def test_server_requests_may_be_released_on_return(emu_lib, tmp_path):
    """A caller may free a request as soon as ikd_server_wait returns it, while
    the worker maintains the request's map: nothing may read it after that."""
    exe = tmp_path / "server_release_check"
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fsanitize=address", "-I", os.path.join(HERE, "..", "include"),
                    os.path.join(HERE, "cpp", "server_release_check.cpp"), "-o", str(exe), "-L", EMU, "-likd_emu",
                    f"-Wl,-rpath,{EMU}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    assert r.stdout.startswith("ok")
