"""Builds libdcpse_hip.so (gfx950) in-tree with hipcc, and the C++ drop-in
test program that links it (tests/cpp/dropin_test)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdcpse_hip.so")
SOURCES = ["dcpse_ctx.hip", "dcpse_support.hip", "dcpse_assemble.hip", "dcpse_apply.hip", "dcpse_sell.hip",
           "dcpse_ode.hip", "dcpse_export.hip", "dcpse_diag.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions stay off: plain float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False, extra=None) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", LIB + ".tmp"]
    if extra:
        cmd[1:1] = list(extra)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
DROPIN_BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


def build_dropin_test(verbose: bool = False) -> str:
    """Host-only C++ program over include/openfpm_dcpse (g++, HIP runtime
    headers, linked against the in-tree libdcpse_hip.so by rpath)."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           DROPIN_SRC, "-o", DROPIN_BIN + ".tmp",
           "-L", HERE, "-ldcpse_hip", "-Wl,-rpath,$ORIGIN/../../openfpm_numerics_amd",
           "-L", os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DROPIN_BIN + ".tmp", DROPIN_BIN)
    return DROPIN_BIN


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
    build_dropin_test(verbose=True)
