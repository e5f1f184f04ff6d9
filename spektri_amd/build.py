"""Build the in-tree HIP library ``spektri_amd/libspektri_gpu.so`` for gfx950.

Usage: ``python -m spektri_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspektri_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 (v_pk_*_f32) is switched off for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function",
]

SOURCES = ["spektri_gpu.hip"]
HOST_DIR = os.path.join(HERE, "host")
HOST_BIN = os.path.join(HOST_DIR, "spektri-gpu")


def build_lib(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    if verbose:
        print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature.
    lines = [ln for ln in (proc.stdout + proc.stderr).splitlines()
             if "packed-fp32-ops" not in ln]
    if lines and verbose:
        print("\n".join(lines), flush=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n" + "\n".join(lines))
    os.replace(tmp, LIB)
    return LIB


HOST_BIN_SAN = os.path.join(HOST_DIR, "spektri-gpu-san")


def build_host(verbose: bool = True) -> str:
    """The spektri-gpu command-line host (host/main.cpp), linked against the
    library with a run path to it, and the same host built with AddressSanitizer
    and UBSan (host code only, for the CPU tests; the library it loads is the
    ordinary build)."""
    for out, extra in ((HOST_BIN, ["-O2"]),
                       (HOST_BIN_SAN, ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined"])):
        tmp = out + ".tmp"
        cmd = [os.environ.get("CXX", "g++"), *extra, "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(HOST_DIR, "main.cpp"), "-L", HERE, "-lspektri_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            raise RuntimeError(f"host build failed ({proc.returncode}):\n" + proc.stdout + proc.stderr)
        os.replace(tmp, out)
    return HOST_BIN


if __name__ == "__main__":
    build_lib()
    build_host()
    sys.exit(0)
