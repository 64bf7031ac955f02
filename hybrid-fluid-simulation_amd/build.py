"""Builds libmgf.so (the HIP library behind include/mgf.h) for gfx950.

Run `python hybrid-fluid-simulation_amd/build.py`; the library is written to
hybrid-fluid-simulation_amd/lib/libmgf.so.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libmgf.so")

# -ffp-contract=off: one rounding per written operation (explicit fmaf where an
# axpy is meant); correctly rounded FP32 divide/sqrt; no packed FP32 ops.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-pthread",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
SOURCES = ["mgf_kernels.hip", "mgf_kernels_mg.hip", "mgf_kernels_ext.hip", "mgf_kernels_particles.hip",
           "mgf_capi.cpp", "mgf_batch.cpp", "mgf_particles.cpp"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_host_test(verbose)
    return LIB


HOST_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "host_classes_test.cpp")
HOST_TEST = os.path.join(LIB_DIR, "host_classes_test")


def build_host_test(verbose=False):
    """The C++ driver of the reference-shaped classes (host/mgf_fluid.hpp):
    host code only, linked against libmgf.so."""
    if not os.path.exists(HOST_TEST_SRC):
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra",
           "-I", os.path.join(os.path.dirname(HERE), "include"), "-I", os.path.join(HERE, "host"),
           HOST_TEST_SRC, "-L", LIB_DIR, "-lmgf", "-Wl,-rpath,$ORIGIN", "-o", HOST_TEST + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(HOST_TEST + ".tmp", HOST_TEST)
    return HOST_TEST


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
