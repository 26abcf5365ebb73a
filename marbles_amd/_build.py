"""Builds libmarbles_amd.so (HIP for gfx950 + C++17 host code) in-tree.

hipcc cross-compiles for gfx950 without a GPU.  Every HIP source is built
with -ffp-contract=off (the FP64 path reproduces the reference arithmetic
bit for bit) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmarbles_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "mrb_level.hip",
    "mrb_compact.hip",
    "mrb_fused.hip",
    "mrb_amr.hip",
    "mrb_batched.hip",
    "mrb_adjoint.hip",
    "mrb_io.hip",
    "mrb_noalloc.hip",
    "mrb_copy.hip",
]

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", INCLUDE,
    "-I", CSRC,
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    err = "\n".join(
        line for line in res.stderr.splitlines()
        if "packed-fp32-ops" not in line and line.strip()
    )
    if err and os.environ.get("MRB_BUILD_VERBOSE"):
        sys.stderr.write(err + "\n")


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_library() -> str:
    """Compiles every HIP source and links libmarbles_amd.so; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_cpp_test() -> str:
    """Compiles tests/cpp/test_cpp_api.cpp (the C++17 host header against the
    library and the CPU oracle) with g++; returns the program's path."""
    src = os.path.join(ROOT, "tests", "cpp", "test_cpp_api.cpp")
    exe = os.path.splitext(src)[0]
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", INCLUDE, "-I", os.path.join(ROOT, "oracle"),
          src, "-o", exe, "-L", HERE, "-lmarbles_amd", "-L", os.path.join(ROOT, "oracle"), "-llbm_oracle",
          "-Wl,-rpath,$ORIGIN/../../marbles_amd", "-Wl,-rpath,$ORIGIN/../../oracle"])
    return exe


if __name__ == "__main__":
    print(build_library())
