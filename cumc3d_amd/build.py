"""Build libcumc3d_hip.so in-tree with hipcc for gfx950, and libcumc3d_io.so
(the HDF5 profile writer, host C) with gcc against the HDF5 C library.

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
# (source, extra flags): the FP32 split's units may contract, use the short
# single-precision divide and square root, and flush single-precision
# denormals (no value of the solver comes near them: the smallest constant
# is WENO-Z's 1e-30), which drops the range-scaling around v_sqrt_f32.  In
# the sweeps' unit (fp32.hip), -fapprox-func with -freciprocal-math also turns
# each divide into v_rcp_f32 and a multiply, without the range check
# (compare, select and two multiplies) the short divide keeps for
# denominators beyond 2^96, and shares the reciprocal of a denominator
# between divides.  The state unit (fp32_state.hip: update, ghost fill,
# FINDDT) keeps IEEE division for its double arithmetic.  The FP64 unit may
# do none of these.
FP32_FLAGS = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]
SOURCES = [(os.path.join(HERE, "csrc", "capi.hip"), []),
           (os.path.join(HERE, "csrc", "fp32.hip"), FP32_FLAGS + ["-fapprox-func", "-freciprocal-math"]),
           (os.path.join(HERE, "csrc", "fp32_state.hip"), FP32_FLAGS)]
DEPS = [os.path.join(HERE, "csrc", f) for f in ("kernels.hpp", "physics.hpp", "selector.hpp", "context.hpp")] + \
       [os.path.join(ROOT, "include", "cumc3d_hip.h")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
IO_LIB = os.path.join(HERE, "libcumc3d_io.so")
IO_SRC = os.path.join(HERE, "csrc", "io.c")
# the HDF5 C library of this image (1.10): headers and libhdf5.so
HDF5_ROOT = os.environ.get("CUMC3D_HDF5", "/opt/conda")
ARCH = os.environ.get("CUMC3D_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add, so the FP64 mode reproduces the
# reference's arithmetic exactly (a later -ffp-contract=fast on the FP32 unit
# overrides it there).  Packed FP32 instructions are disabled.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-rdc",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{os.path.basename(cmd[0])} failed: " + " ".join(cmd[-3:]))


def build_io(verbose: bool = True) -> str | None:
    """libcumc3d_io.so; skipped (None) where no HDF5 C library is installed."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"no HDF5 headers under {inc}: libcumc3d_io.so not built")
        return None
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-fPIC", "-shared", f"-I{inc}", IO_SRC, "-o", IO_LIB + ".tmp",
          f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}", "-lm"])
    os.replace(IO_LIB + ".tmp", IO_LIB)
    if verbose:
        print(f"built {IO_LIB}")
    return IO_LIB


def build(verbose: bool = True) -> str:
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    build_io(verbose)
    objs = [os.path.join(HERE, "csrc", os.path.basename(src) + ".o") for src, _ in SOURCES]
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        list(ex.map(_run, [[HIPCC, *FLAGS, *extra, "-c", src, "-o", obj]
                           for (src, extra), obj in zip(SOURCES, objs)]))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-rdc", "-o", LIB + ".tmp", *objs])
    for o in objs:
        os.remove(o)
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s")
    return LIB


if __name__ == "__main__":
    build()
