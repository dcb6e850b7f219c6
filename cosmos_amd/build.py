"""Builds the gfx950 HIP library of the finite-volume update in-tree.

    python -m cosmos_amd.build            # libcosmos_fv.so next to this file

The library is plain hipcc output (no torch extension): the C-ABI in
include/cosmos_fv.h over the kernels in csrc/.  FP64 code is compiled with
-ffp-contract=off (hipcc contracts by default for HIP) so that the only fused
multiply-adds are the explicit ones, and without fast-math, so divisions and
square roots are correctly rounded.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcosmos_fv.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("COSMOS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["fv_kernels.hip", "fv_gr.hip", "fv_capi.hip", "fv_amr.hip", "fv_adjoint.hip", "fv_gr_adjoint.hip",
           "fv_batch.hip", "fv_pointwise.hip", "fv_noalloc.hip", "fv_point.hip"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


ROOT = os.path.dirname(HERE)
DROPIN = os.path.join(HERE, "dropin")
DROPIN_LIB = os.path.join(HERE, "libcosmos_dropin.so")
DROPIN_CHECK = os.path.join(ROOT, "tests", "dropin", "dropin_check")
CXX = os.environ.get("CXX", "g++")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> str:
    """Compile the HIP library (libcosmos_fv.so), the C++ drop-in classes over
    it (libcosmos_dropin.so) and the drop-in check program; returns the HIP
    library path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *FLAGS, "-o", tmp, *srcs], verbose)
    os.replace(tmp, LIB)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + DROPIN, "-I" + os.path.join(DROPIN, "compat")]
    _run([CXX, "-std=c++14", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", *inc, "-o", DROPIN_LIB,
          os.path.join(DROPIN, "cosmos_dropin.cpp"), os.path.join(DROPIN, "cosmos_dropin_point.cpp"), "-L" + HERE,
          "-lcosmos_fv", "-Wl,-rpath,$ORIGIN"], verbose)
    _run([CXX, "-std=c++14", "-O2", "-Wall", *inc, "-o", DROPIN_CHECK, DROPIN_CHECK + ".cpp", "-L" + HERE,
          "-lcosmos_dropin", "-lcosmos_fv", "-Wl,-rpath," + HERE], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
