"""Builds the gfx950 engine library (libpmx.so) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libpmx.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("pmx.hip", "pmx_kernels.h", "pmx_batch.h", "pmx_common.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 arithmetic is kept off (only scalar/FMA FP32 VALU forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build_lib(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["hipcc", *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "csrc", "pmx.hip"), "-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


HOST_EXE = os.path.join(HERE, "lib", "pmx_force")


def build_host(verbose: bool = False) -> str:
    """The C++ ProtoMol-boundary layer and its driver, linked to libpmx.so."""
    host = os.path.join(HERE, "host")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(host, "pmx_protomol.cpp"), os.path.join(host, "pmx_force.cpp"),
           "-L" + os.path.dirname(LIB), "-lpmx", "-Wl,-rpath,$ORIGIN", "-o", HOST_EXE + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stderr[-4000:])
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    if verbose:
        print("built", HOST_EXE)
    return HOST_EXE


SERVE_LIB = os.path.join(HERE, "lib", "libpmx_serve.so")


def build_serve(verbose: bool = False) -> str:
    """The request loop (host C++ over the C ABI), linked to libpmx.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-pthread", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "host", "pmx_serve.cpp"), "-L" + os.path.dirname(LIB), "-lpmx",
           "-Wl,-rpath,$ORIGIN", "-o", SERVE_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stderr[-4000:])
    os.replace(SERVE_LIB + ".tmp", SERVE_LIB)
    if verbose:
        print("built", SERVE_LIB)
    return SERVE_LIB


if __name__ == "__main__":
    build_lib(verbose=True)
    build_host(verbose=True)
    build_serve(verbose=True)
