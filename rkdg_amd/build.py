"""Builds the HIP library rkdg_amd/librkdg.so for gfx950 (in-tree, no JIT cache).

hipcc compiles the kernel translation units and the host C++ that implements the
C-ABI of include/rkdg_capi.h (one object per source, in parallel), then links
them; the library has no torch dependency.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librkdg.so")
SOURCES = ["rkdg_tensor.hip", "rkdg_dense.hip", "rkdg_batch.hip", "rkdg_adjoint.hip", "rkdg_noalloc.hip", "rkdg_io.hip", "rkdg_shu.hip",
           "rkdg_nc.hip", "rkdg_host.cpp", "rkdg_selftest.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{os.path.basename(cmd[0])} failed: " + " ".join(cmd[-3:]))


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_callers(verbose)
    return LIB


def build_callers(verbose: bool = False) -> None:
    """The compiled callers of the C-ABI that the tests run: a C program on the
    plain header (gcc) and a C++ program on the class mirror include/rkdg.hpp
    (hipcc for the HIP runtime's allocation calls); both link librkdg.so."""
    root = os.path.dirname(HERE)
    capi = os.path.join(root, "tests", "capi")
    if not os.path.isdir(capi):
        return
    inc = ["-I", os.path.join(root, "include"), "-I", capi]
    link = ["-L", HERE, "-lrkdg", "-Wl,-rpath,$ORIGIN/../../rkdg_amd"]
    _run(["gcc", "-std=c11", "-O2", "-Wall", *inc, os.path.join(capi, "capi_smoke.c"), *link, "-lm",
          "-o", os.path.join(capi, "capi_smoke")], verbose)
    _run([HIPCC, *FLAGS, "-Wall", *inc, os.path.join(capi, "cpp_smoke.cpp"), *link,
          "-o", os.path.join(capi, "cpp_smoke")], verbose)


if __name__ == "__main__":
    print(build(verbose=True))
