"""Build libfsgh.so (HIP, gfx950) in-tree with hipcc.

The library is compiled without floating-point contraction (-ffp-contract=off)
so that every FP64 expression is evaluated in the written operation order,
matching the reference arithmetic.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfsgh.so")
SOURCES = ["fsgh_kernels.hip", "fsgh_mesh.hip", "fsgh_api.hip", "fsgh_adjoint.hip",
           "fsgh_setup.hip", "fsgh_tune.hip", "fsgh_rz.hip", "fsgh_stage.hip"]
ARCH = os.environ.get("FSGH_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into fierro_amd/libfsgh.so and return its path."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
