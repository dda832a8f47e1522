"""The C++ mirror of the reference interface (include/mdy_medyan.hpp), built
with g++ against libmdy_hip.so and run as a separate process."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_medyan_shim.cpp")
LIBDIR = os.path.join(ROOT, "medyan_amd")


def build(d):
    exe = os.path.join(d, "test_medyan_shim")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe,
                    "-L", LIBDIR, "-lmdy_hip", f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_shim_compiles():
    with tempfile.TemporaryDirectory() as d:
        assert os.path.exists(build(d))


@pytest.mark.gpu
def test_shim_runs_on_device():
    with tempfile.TemporaryDirectory() as d:
        exe = build(d)
        print("[gpu] running the C++ interface test", flush=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0 and "ALL OK" in r.stdout
