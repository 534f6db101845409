"""Builds libfluidsim_hip.so in-tree with hipcc for gfx950.

Run as ``python -m fluidsim_amd.build`` (or through ``__graft_entry__.build()``).  The library is
written next to this file so it travels with the source tree; nothing is JIT-compiled at import.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfluidsim_hip.so")
ARCH = os.environ.get("FLUIDSIM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["fs_kernels.hip", "fs_multigrid.hip", "fs_transport.hip", "fs_viscosity.hip", "fs_tension.hip", "fs_host.cpp",
           "fs_server.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # FP64 coefficient and hierarchy kernels must round exactly like the reference's C++.
    "-ffp-contract=off",
    # No packed FP32 arithmetic in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cc = hipcc()

    def compile_one(src):
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *HIPCC_FLAGS, *lang, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        return src, obj, subprocess.run(cmd, capture_output=True, text=True)

    # one hipcc per source, a few at a time (each is single-threaded)
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "4")), os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        results = list(pool.map(compile_one, SOURCES))
    objs = []
    for src, obj, proc in results:
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    build_cpp_test(verbose)
    return LIB


ROOT = os.path.dirname(HERE)
CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "grid_api_test.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "grid_api_test")


def build_cpp_test(verbose: bool = False) -> str:
    """The C++ adapter test (host code only): g++ against the C ABI library."""
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT}", "-I/opt/rocm/include", CPP_TEST_SRC,
           f"-L{HERE}", "-lfluidsim_hip", "-Wl,-rpath,$ORIGIN/../../fluidsim_amd", "-o", CPP_TEST_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("g++ failed on the C++ adapter test")
    return CPP_TEST_BIN


if __name__ == "__main__":
    print(build(verbose=True))
