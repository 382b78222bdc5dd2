"""Build ``libcsg.so`` (HIP kernels + C ABI + host RNG) in-tree for gfx950.

Usage: ``python build.py`` (or ``__graft_entry__.build()``).  Compiles every
time it is called.  Objects go to ``_build/`` next to this file; the shared
library is written beside this file so it travels with the tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcsg.so")
ARCH = os.environ.get("CSG_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 VALU ops are disabled: hipcc -O3 would otherwise form them from
# ordinary float code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]

# The kernel instantiations are split by (precision, size family) so the
# translation units compile in parallel.
UNITS = [
    ("csg_capi.hip", []),
    ("csg_batch.hip", []),
    ("csg_noalloc.hip", []),
    ("csg_persist.hip", []),
] + [(f"csg_ops_{prec}_{group}.hip", []) for prec in ("fp64", "mixed") for group in ("small", "mid", "big", "pow3")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    rng_obj = os.path.join(BUILD, "julia_rng.o")
    jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "julia_rng.cpp"), "-o", rng_obj])
    objs.append(rng_obj)
    with ThreadPoolExecutor(max_workers=min(int(os.environ.get("CSG_BUILD_JOBS", "8")), len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
