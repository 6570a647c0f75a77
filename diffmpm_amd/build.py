"""Builds the HIP library libdiffmpm_amd.so in-tree for gfx950.

``python -m diffmpm_amd.build`` or ``__graft_entry__.build()``.  The library
is a plain C-ABI shared object (``include/diffmpm_amd.h``) loaded by ctypes.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdiffmpm_amd.so")
SOURCES = ["mpm_forward.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # a * b + c within one source expression becomes one fused multiply-add
    # (llvm.fmuladd), never across statements
    "-ffp-contract=on",
    # Keep the compiler from emitting packed FP32 (v_pk_*_f32) instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    extra = [f for f in sorted(os.listdir(CSRC)) if f.endswith(".hip") and f not in SOURCES]
    return [os.path.join(CSRC, f) for f in SOURCES + extra]


def build_library(verbose=False):
    """Compiles every .hip source under csrc/ and links the shared library."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in sources():
        obj = os.path.join(build_dir, os.path.basename(src) + ".o")
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
