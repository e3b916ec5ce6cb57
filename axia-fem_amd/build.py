"""Build libaxia_gpu.so for gfx950 with hipcc (no GPU needed to compile).

Usage: python axia-fem_amd/build.py [--jobs N]
Output: axia-fem_amd/lib/libaxia_gpu.so
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build", "obj")
LIB = os.path.join(HERE, "lib", "libaxia_gpu.so")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms stay off (hipcc emits them from plain float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]
# Files whose FP64 results must follow the reference's operation order.
EXACT = {"axg_element.hip", "axg_triplets.hip", "axg_loads.cpp", "axg_deck.cpp", "axg_writers.cpp"}
SOURCES = [
    "axg_api.cpp", "axg_symbolic.cpp", "axg_loads.cpp", "axg_element.hip", "axg_triplets.hip",
    "axg_pcg.hip", "axg_mixed.hip", "axg_element32.hip", "axg_bench.hip", "axg_batch.hip",
    "axg_deck.cpp", "axg_eigen.hip", "axg_writers.cpp", "axg_noalloc.hip",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def compile_one(src, obj_dir=OBJ, extra=()):
    obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    flags = list(COMMON) + list(extra)
    if src in EXACT:
        flags.append("-ffp-contract=off")
    if src.endswith(".cpp"):
        flags = ["-x", "hip"] + flags
    cmd = [hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True, lib=LIB, obj_dir=OBJ, extra=(), link_extra=()):
    """Compile every source and link lib. extra / link_extra: additional compile / link
    flags (the sanitizer builds of sanitize.py use them)."""
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(os.path.dirname(lib), exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: compile_one(src, obj_dir, extra), SOURCES))
    tmp = lib + ".tmp"
    cmd = ([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs +
           ["-lpthread"] + list(link_extra))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
