"""The C++ mirror of the reference interface (include/mdy_medyan.hpp), built
with g++ against libmdy_hip.so as a shared object and run inside the test
process (one process drives the GPU)."""
import ctypes
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_medyan_shim.cpp")
LIBDIR = os.path.join(ROOT, "medyan_amd")


def build(d):
    so = os.path.join(d, "libtest_medyan_shim.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"), SRC,
                    "-o", so, "-L", LIBDIR, "-lmdy_hip", f"-Wl,-rpath,{LIBDIR}"], check=True)
    return so


def test_shim_compiles():
    with tempfile.TemporaryDirectory() as d:
        assert os.path.exists(build(d))


@pytest.mark.gpu
def test_shim_runs_on_device():
    with tempfile.TemporaryDirectory() as d:
        lib = ctypes.CDLL(build(d))
        print("[gpu] running the C++ interface test", flush=True)
        lib.mdy_shim_run.restype = ctypes.c_int
        rc = lib.mdy_shim_run()
        assert rc == 0, f"C++ interface check failed (line {rc})"
