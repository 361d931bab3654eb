"""Builds the native libraries in-tree.

- ``steadyscan_amd/libsteadyscan_amd.so``: HIP kernels + C ABI
  (include/steadyscan.h), compiled for gfx950 with hipcc.
- ``steadyscan_amd/libss_synth.so``: the seeded synthetic-data generator
  (include/steadyscan_synth.h), host C++.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

LIB = os.path.join(HERE, "libsteadyscan_amd.so")
SYNTH_LIB = os.path.join(HERE, "libss_synth.so")

HIP_SOURCES = ["ss_kernels.hip", "ss_engine.hip", "ss_resample.hip", "ss_ibs.hip", "ss_prefilter.hip"]
CXX_SOURCES = ["ss_geometry.cpp", "ss_driver.cpp", "ss_rom.cpp"]

# packed FP32 instructions are disabled (they have produced silent errors on
# this hardware pool); hipcc's host pass warns about the feature, harmlessly.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    return r


def build_synth(verbose=False):
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-I", INCLUDE,
          os.path.join(CSRC, "synth", "ss_synth.cpp"), "-o", SYNTH_LIB + ".tmp"], verbose)
    os.replace(SYNTH_LIB + ".tmp", SYNTH_LIB)


def build_hip(verbose=False):
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        _run([hipcc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        # host C++ that shares __host__ __device__ formulas with the kernels
        # (ss_scan.h) sees the HIP headers in host-only form
        _run([shutil.which("g++") or "g++", "-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
              "-I", os.path.join(ROCM, "include"), "-I", INCLUDE, "-I", CSRC, "-c",
              os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)


def build(verbose=False):
    build_synth(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, SYNTH_LIB)
