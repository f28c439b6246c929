"""Builds libehip.so (HIP for gfx950) and the Fortran interface module in-tree.

    python -m elmerfem_amd.build          # compile everything
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libehip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_SOURCES = ["system.hip", "krylov.hip", "assembly.hip", "batch.hip", "noalloc.hip", "ilu.hip", "pattern_gpu.hip"]
CXX_SOURCES = ["pattern.cpp", "meshio.cpp"]

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_lib(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fopenmp", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose=True):
    """Compiles the ISO_C_BINDING module and the example driver (host only)."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found; skipping Fortran interface")
        return None
    out = os.path.join(fdir, "ehip_demo")
    _run([FLANG, "-O2", "-J", OBJ, "-o", out, os.path.join(fdir, "ehip.F90"), os.path.join(fdir, "ehip_demo.F90"),
          "-L" + HERE, "-lehip", "-Wl,-rpath," + HERE])
    if verbose:
        print("built", out)
    return out


def main():
    build_lib()
    if os.path.exists(os.path.join(HERE, "fortran", "ehip.F90")):
        build_fortran()


if __name__ == "__main__":
    sys.exit(main())
