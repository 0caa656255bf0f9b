"""Build libfsgh.so (HIP, gfx950) in-tree with hipcc.

The library is compiled without floating-point contraction (-ffp-contract=off)
so that every FP64 expression is evaluated in the written operation order,
matching the reference arithmetic.  Each source is compiled to an object in
parallel (no relocatable device code: every translation unit carries its own
kernels), then the objects are linked into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfsgh.so")
SOURCES = ["fsgh_kernels.hip", "fsgh_mesh.hip", "fsgh_api.hip", "fsgh_adjoint.hip",
           "fsgh_setup.hip", "fsgh_tune.hip", "fsgh_rz.hip", "fsgh_stage.hip",
           "fsgh_noalloc.hip"]
ARCH = os.environ.get("FSGH_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
]


def _jobs() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16, len(SOURCES)))


def build(verbose: bool = False, out: str = LIB, defines=(), sources=None) -> str:
    """Compile every HIP source into ``out`` (default fierro_amd/libfsgh.so)
    and return its path.  ``defines`` adds -D options to every source (used
    by test-only variants of the library)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in (sources or SOURCES)]
    dflags = [f"-D{d}" for d in defines]
    with tempfile.TemporaryDirectory(prefix="fsgh_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.basename(s) + ".o") for s in srcs]

        def compile_one(pair):
            src, obj = pair
            cmd = [hipcc, *FLAGS, *dflags, "-c", "-o", obj, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            return src, r

        with ThreadPoolExecutor(_jobs()) as ex:
            results = list(ex.map(compile_one, zip(srcs, objs)))
        for src, r in results:
            err = "\n".join(l for l in r.stderr.splitlines()
                            if "packed-fp32-ops" not in l)
            if err.strip():
                print(err, file=sys.stderr, flush=True)
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, f"hipcc -c {src}")
        tmp = out + ".tmp"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
    return out


STRENGTH_TEST_LIB = os.path.join(HERE, "libfsgh_strength_test.so")


def build_test_variants(verbose: bool = False) -> list:
    """Test-only builds of the same sources: libfsgh_strength_test.so, whose
    user strength hook runs the hypoelastic test model (the product library
    ships the reference's empty body)."""
    return [build(verbose, out=STRENGTH_TEST_LIB, defines=("FSGH_TEST_STRENGTH_MODEL",))]


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_test_variants(verbose="-v" in sys.argv)
