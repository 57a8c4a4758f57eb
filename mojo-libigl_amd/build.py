"""Build the MI355X winding-number library (gfx950 HIP + host C++).

    python mojo-libigl_amd/build.py        # -> mojo-libigl_amd/lib/libmojo_libigl_amd.so

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled (-packed-fp32-ops); the CPU path is compiled by g++ for
x86-64-v3 with FMA contraction off. No incremental logic: every call rebuilds.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmojo_libigl_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_SOURCES = ["wn_forward.hip", "wn_strip.hip", "wn_strip64.hip", "wn_sort.hip", "wn_batch.hip", "wn_backward.hip",
               "wn_aabb.hip", "wn_fwn.hip", "wn_ribbon.hip", "wn_capi.hip"]
CPU_SOURCES = ["wn_cpu.cpp", "wn_strips.cpp", "wn_edges.cpp", "wn_aabb_tree.cpp", "wn_fwn_tree.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", CSRC,
    "-I", INCLUDE,
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v3", "-ffp-contract=off",
             "-fno-fast-math", "-Wall", "-Wextra", "-I", INCLUDE]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + proc.stdout + proc.stderr)


def _compile(src: str, extra: list[str]) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        _run([HIPCC, *HIP_FLAGS, *extra, "-c", path, "-o", obj])
    else:
        _run(["g++", *CXX_FLAGS, "-c", path, "-o", obj])
    return obj


def build(extra_hip_flags: list[str] | None = None, jobs: int = 8) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    sources = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))] + CPU_SOURCES
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as pool:
        objs = list(pool.map(lambda s: _compile(s, extra_hip_flags or []), sources))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(sys.argv[1:]))
