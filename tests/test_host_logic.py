"""Configuration errors of the C++ host layer (mara2_amd/host/mara2.hpp) that
need no GPU: compiled with g++ against the header and the built libraries,
run on the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mara2_amd")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_layer_configuration_errors(tmp_path):
    for lib in ("libmara_gpu.so", "libmara2_host.so"):
        if not os.path.exists(os.path.join(PKG, lib)):
            pytest.fail(f"{lib} is not built (run __graft_entry__.build())")
    exe = tmp_path / "logic_errors"
    src = os.path.join(ROOT, "tests", "host", "logic_errors.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), src, f"-L{PKG}", "-lmara2_host", "-lmara_gpu",
                    f"-Wl,-rpath,{PKG}"], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 6, r.stdout
