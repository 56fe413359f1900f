"""Compile flatworld_amd/libflatworld_fem.so for gfx950 with hipcc.

Floating-point flags are part of the numerics contract: ``-ffp-contract=off``
keeps every multiply and add separately rounded (the CPU oracle is built the
same way), HIP's default correctly-rounded float32 division and square root
stay on, and packed-fp32 instructions are disabled.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflatworld_fem.so")
SOURCES = ["fw_kernels.hip", "fw_flex.hip", "fw_arena.hip"]
ARCH = os.environ.get("FW_OFFLOAD_ARCH", "gfx950")


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; install ROCm or put hipcc on PATH")
    return exe


def compile_flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-result",
        f"-I{os.path.join(ROOT, 'include')}",
    ]


def build(verbose=False):
    """Compile every HIP source and link the shared library in-tree."""
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [hipcc(), *compile_flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
