"""Builds libmdy_hip.so in-tree with hipcc for gfx950.

Every HIP source is compiled with -ffp-contract=off (the FP64 path reproduces
the CPU restatement's rounding) and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmdy_hip.so")
TRAJ_OUT = os.path.join(HERE, "libmdy_traj.so")
# the HDF5 C library for the traj.h5 writer (a separate library, optional)
HDF5_ROOT = os.environ.get("MDY_HDF5_ROOT", "/opt/conda")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["mdy_api.hip", "mdy_eval.hip", "mdy_nl.hip", "mdy_cg.hip", "mdy_hvp.hip", "mdy_aux.hip",
               "mdy_quad.hip"]
CXX_SOURCES = ["netgen.cpp"]
ARCH = os.environ.get("MDY_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
HIPFLAGS = [f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["hipcc", "-c", *HIPFLAGS, *COMMON, os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++", "-c", *COMMON, "-ffp-contract=off", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = OUT + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, OUT)
    build_traj()
    return OUT


def build_traj():
    """libmdy_traj.so (traj.h5 writer) when the HDF5 headers are present;
    returns its path or None."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        return None
    tmp = TRAJ_OUT + ".tmp"
    _run(["g++", *COMMON, "-shared", "-I", inc, os.path.join(CSRC, "traj_h5.cpp"), "-o", tmp,
          "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}"])
    os.replace(tmp, TRAJ_OUT)
    return TRAJ_OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
