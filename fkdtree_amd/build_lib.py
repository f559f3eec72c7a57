"""Build recipe of the native library (gfx950 HIP, one translation unit).

``python -m fkdtree_amd.build_lib`` compiles ``csrc/fkd_capi.hip`` into
``fkdtree_amd/libfkdtree_hip.so`` in place.  hipcc cross-compiles for gfx950
without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfkdtree_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "fkd_capi.hip")]
ARCH = os.environ.get("FKD_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def hipcc_cmd(out: str = LIB, extra=()) -> list:
    return [
        HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # Packed FP32 instructions stay disabled for every kernel.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
        "-Wl,-rpath,/opt/rocm/lib",
        *extra, *SOURCES, "-o", out,
    ]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = hipcc_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose and res.stderr.strip():
        print(res.stderr)
    os.replace(tmp, LIB)
    return LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "fkdtree_hip_test.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "fkdtree_hip_test.bin")
CPP_PERBOX_SRC = os.path.join(ROOT, "tests", "cpp", "fkdtree_perbox.cpp")
CPP_PERBOX_BIN = os.path.join(ROOT, "tests", "cpp", "fkdtree_perbox.bin")


def build_cpp_test(verbose: bool = False) -> str:
    """The C++ drop-in drivers (host code only: g++ against the C ABI): the
    parity driver and the per-box search driver."""
    for src, out in ((CPP_TEST_SRC, CPP_TEST_BIN), (CPP_PERBOX_SRC, CPP_PERBOX_BIN)):
        cmd = ["g++", "-std=c++14", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), src,
               "-L", HERE, "-lfkdtree_hip", "-Wl,-rpath,$ORIGIN/../../fkdtree_amd", "-pthread", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("g++ failed:\n" + res.stdout + res.stderr)
    return CPP_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
