"""Compile the HIP library (gfx950) in-tree.

``python -m galacticus_amd.build`` produces ``galacticus_amd/lib/libgalacticus_amd.so``.
The library is plain C ABI (include/gal_ode.h); nothing here depends on torch.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libgalacticus_amd.so")
SOURCES = ["gal_ode.hip", "surrogate.hip", "tree_exec.hip", "trees.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops off: hipcc must not emit packed FP32 instructions.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-result"]
# Per-source flags.  The integrator kernel is one loop around the whole RHS;
# machine-level loop-invariant code motion hoists the RHS's 64-bit constants
# (polynomial coefficients of exp/log and the like) out of it and holds them
# in registers for the whole integration, which spills live values to
# scratch memory.  Materialising a constant where it is used costs the same
# two moves and no register.
SOURCE_FLAGS = {"gal_ode.hip": ["-mllvm", "-disable-machine-licm"]}


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs, objs = [], []
    for s in srcs:
        obj = os.path.join(objdir, s + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *SOURCE_FLAGS.get(s, []), "-c", "-o", obj, os.path.join(CSRC, s)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(s)
    if failed:
        raise RuntimeError("hipcc failed compiling " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libgalacticus_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
