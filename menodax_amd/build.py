"""Build the native part in-tree: ``libmenodax_hip.so``, the C ABI of
``include/menodax_hip.h``. (An XLA FFI handler over that ABI would be built the
same way against jaxlib's FFI headers; none is shipped, since no jaxlib exists
where this library is built and tested.)

The per-right-hand-side kernels are not built here: they are generated and
compiled with hipRTC when a solve first needs them (``menodax_amd._runtime``).
``python -m menodax_amd.build`` runs this.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmenodax_hip.so")

# Every HIP translation unit is built with packed-FP32 instructions disabled.
HIP_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; set HIPCC or install ROCm")


def build_runtime(verbose: bool = False) -> str:
    cmd = [
        _hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
        *HIP_FLAGS,
        "-I", os.path.join(ROOT, "include"),
        os.path.join(HERE, "csrc", "runtime.cpp"),
        "-o", LIB + ".tmp",
        "-lhiprtc",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_all(verbose: bool = False):
    build_runtime(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
