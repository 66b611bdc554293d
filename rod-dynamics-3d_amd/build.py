"""Builds librdk.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled with -ffp-contract=off (separately rounded
multiplies and adds, so the FP64 narrow phase and integrators round exactly
like the sequential CPU evaluation of the same expressions) and with packed
FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librdk.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RDK_OFFLOAD_ARCH", "gfx950")

SOURCES = ["rdk_api.hip", "rdk_detect.hip", "rdk_nsc.hip", "rdk_forces.hip",
           "rdk_integrate.hip", "rdk_sort.hip", "rdk_mlp.hip", "rdk_entangle.hip", "rdk_populate.hip",
           "rdk_gen.cpp"]

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = []
    for src, obj, rc, out in results:
        if rc != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{out}")
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_facade_check()
    return LIB


FACADE_SRC = os.path.join(HERE, "cpp", "facade_check.cpp")
FACADE_LIB = os.path.join(HERE, "librdk_facade_check.so")


def build_facade_check():
    """The C++ classes of include/rdk_physics.hpp behind two C entry points
    (librdk_facade_check.so, host code only, linked against librdk.so)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = FACADE_LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-I", inc, FACADE_SRC,
           "-L", HERE, "-lrdk", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade check build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, FACADE_LIB)
    return FACADE_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
