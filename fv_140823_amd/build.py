"""Build the native library of the package and the CPU oracle.

libfv140823.so = host C (forest, flattener, RUN binding) + HIP kernels for
gfx950, linked by hipcc.  The oracle (test infrastructure) is built by its own
Makefile under oracle/.  Both are compiled in-tree so a copy of the tree carries
the built libraries.
"""
from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "_lib" / "libfv140823.so"
ORACLE_DIR = REPO / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HOST_SOURCES = ["host/params.c", "host/forest.c", "host/flatten.c", "host/bind.c", "host/export.c"]
HIP_SOURCES = ["fvgpu.hip", "fvadj.hip", "fvmesh.hip", "fvtile.hip"]

ARCH = os.environ.get("FV_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions stay off (their forms have misbehaved on this hardware)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CC_FLAGS = ["-O2", "-fPIC", "-std=c11", "-pthread", "-Wall", "-Wno-unused-function"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP kernels cannot be built")


def build_library(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    inc = [f"-I{REPO / 'include'}", f"-I{CSRC}", f"-I{CSRC / 'host'}"]
    objs = []
    procs = []
    for src in HOST_SOURCES:
        obj = BUILD / (pathlib.Path(src).stem + ".o")
        cmd = ["gcc", *CC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = BUILD / (pathlib.Path(src).stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if verbose:
            print("+", " ".join(cmd), flush=True)
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed = True
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> pathlib.Path:
    _run(["make", "-s", "-B", "-C", str(ORACLE_DIR)], verbose)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB)
    print(ORACLE_LIB)
