"""Build recipe for libcsu.so (host C + gfx950 HIP kernels) and the drop-in
``executable``.  Everything lands in ``csu_amd/lib/`` (in-tree, so the built
files travel with the source tree).

    python -m csu_amd.build            # build everything
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HOST_SRC = ["input.c", "mesh.c", "case.c", "output.c", "order.c"]
HIP_SRC = ["csu_dev.hip"]
CXX_SRC = ["csu_serve.cpp"]   # host C++ over the C ABI (no device code)
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: the FP64 path must round every product as the
# reference's x86-64 build does (no fused multiply-add).
# -packed-fp32-ops off: no v_pk_*_f32 in any kernel.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + os.path.join(PKG, "csrc")]
HOST_FLAGS = ["-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-std=gnu11", "-I" + INC]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    elif verbose and (r.stdout or r.stderr):
        out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
        if out.strip():
            sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> str:
    """Compile and link; returns the path of libcsu.so."""
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(LIB, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for s in HOST_SRC:
        o = os.path.join(objdir, s.replace(".c", ".o"))
        _run([CC, *HOST_FLAGS, "-c", os.path.join(PKG, "host", s), "-o", o], verbose)
        objs.append(o)
    for s in CXX_SRC:
        o = os.path.join(objdir, s.replace(".cpp", ".o"))
        _run([CXX, "-O2", "-fPIC", "-std=c++17", "-Wall", "-I" + INC, "-c", os.path.join(PKG, "csrc", s), "-o", o],
             verbose)
        objs.append(o)
    for s in HIP_SRC:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(PKG, "csrc", s), "-o", o], verbose)
        objs.append(o)
    so = os.path.join(LIB, "libcsu.so")
    tmp = so + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lm"], verbose)
    os.replace(tmp, so)
    exe = os.path.join(LIB, "executable")
    _run([CC, *HOST_FLAGS, os.path.join(PKG, "host", "main.c"), "-o", exe + ".tmp",
          "-L" + LIB, "-lcsu", "-Wl,-rpath,$ORIGIN", "-lm"], verbose)
    os.replace(exe + ".tmp", exe)
    shutil.rmtree(objdir, ignore_errors=True)
    return so


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
