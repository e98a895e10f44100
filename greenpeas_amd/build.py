"""Builds the gfx950 library libgreenpeas_amd.so in-tree with hipcc.

Usage: python -m greenpeas_amd.build   (or greenpeas_amd.build.build())
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgreenpeas_amd.so")
SOURCES = ["circuit.cpp", "workers.cpp", "stepg.cpp", "engine.hip", "api.cpp"]
ARCH = os.environ.get("GREENPEAS_AMD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compiles every source for gfx950 into one shared library."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    common = [
        "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include"),
    ]
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *common, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_pymodule(verbose)
    build_facade_check(verbose)
    return LIB


def build_pymodule(verbose: bool = False) -> str:
    """The pybind11 module greenpeas_amd._gppy (host code over the C++
    facade, linked to libgreenpeas_amd.so next to it)."""
    import sysconfig

    import pybind11

    out = os.path.join(HERE, "_gppy" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           os.path.join(CSRC, "gppy.cpp"), "-o", tmp, "-L", HERE, "-l:libgreenpeas_amd.so",
           "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_facade_check(verbose: bool = False) -> str:
    """tests/cpp/facade_check: a C++ program over include/greenpeas_amd.hpp
    that the tests run (it compiles a circuit file and prints the DEM)."""
    src = os.path.join(HERE, "..", "tests", "cpp", "facade_check.cpp")
    out = os.path.join(HERE, "..", "tests", "cpp", "facade_check")
    if not os.path.exists(src):
        return ""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"), src, "-o", out + ".tmp",
           "-L", HERE, "-l:libgreenpeas_amd.so", "-Wl,-rpath,$ORIGIN/../../greenpeas_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
