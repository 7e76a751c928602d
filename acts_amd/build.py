"""Build recipe for the native library libacts_kf.so (HIP, gfx950).

Run as ``python -m acts_amd.build`` or call :func:`build_library`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libacts_kf.so")

SOURCES = ["kf_capi.hip", "generator.hip"]
HEADERS = ["kf_common.hpp", "kf_device.hpp", "kf_mixed.hpp", "kf_fit.hpp", "kf_kernel.hpp",
           "kf_reference.hpp", "kf_reverse.hpp", "kf_states.hpp", "kf_field_host.hpp"]

# -ffp-contract=off: the FP64 fit keeps the reference's separate multiply and
# add roundings (x86-64 without -march has no FMA), so every decision of the
# adaptive stepper sees the same numbers as the CPU fitter.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build_library(verbose: bool = False, extra_flags: list[str] | None = None) -> str:
    """Compile libacts_kf.so in-tree; returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, *(extra_flags or []), "-o", tmp, *srcs,
           "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libacts_kf.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


FACADE_SRC = os.path.join(REPO_DIR, "tests", "cpp", "facade_check.cpp")
FACADE_BIN = os.path.join(REPO_DIR, "tests", "cpp", "facade_check")


def build_facade_check(verbose: bool = False) -> str:
    """Compile the C++ facade example against the built library (host only)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO_DIR, "include"),
           "-o", FACADE_BIN + ".tmp", FACADE_SRC, "-L", PKG_DIR, "-lacts_kf",
           "-Wl,-rpath," + PKG_DIR, "-Wl,-rpath,$ORIGIN/../../acts_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building facade_check")
    os.replace(FACADE_BIN + ".tmp", FACADE_BIN)
    return FACADE_BIN


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_facade_check(verbose=True))
