"""Build libcumc3d_hip.so in-tree with hipcc for gfx950.

usage: python -m cumc3d_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcumc3d_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "capi.hip")]
DEPS = [os.path.join(HERE, "csrc", f) for f in ("kernels.hpp", "physics.hpp")] + \
       [os.path.join(ROOT, "include", "cumc3d_hip.h")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CUMC3D_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add, so the FP64 mode reproduces the
# reference's arithmetic exactly.  Packed FP32 instructions are disabled.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-rdc",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = True) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build()
