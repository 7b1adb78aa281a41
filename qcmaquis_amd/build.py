"""Build the gfx950 product library in-tree.

``python -m qcmaquis_amd.build`` compiles every source with hipcc for gfx950 and links
``qcmaquis_amd/libqcmaquis_amd.so``.  It always recompiles (no dependency tracking).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libqcmaquis_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QCMAQUIS_AMD_ARCH", "gfx950")

# packed FP32 instructions are disabled for all device code
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + os.path.basename(cmd[-1]))
    return r


def build_library(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    sources = sorted(f for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))
    jobs = []
    for src in sources:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        # host-only translation units: plain C++ against the HIP runtime headers
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-c",
                                  "-o", obj]
        if src.endswith(".hip"):
            cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", "-o", obj]
        jobs.append((cmd + [os.path.join(CSRC, src)], obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-fopenmp", "-o", tmp] + [o for _, o in jobs] +
         ["-L/opt/rocm/lib", "-lrocblas", "-lrocsolver"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


H5_LIB = os.path.join(PKG, "libqcmaquis_amd_h5.so")
H5_PREFIX = os.environ.get("QCMAQUIS_AMD_HDF5", "/opt/conda")


def build_h5(verbose=False):
    """the host-only result / checkpoint file library (HDF5 C library under H5_PREFIX); skipped,
    returning None, when the HDF5 headers are absent"""
    inc, lib = os.path.join(H5_PREFIX, "include"), os.path.join(H5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print("HDF5 headers not found under", H5_PREFIX, "- result files not built")
        return None
    tmp = H5_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", inc,
          "-o", tmp, os.path.join(CSRC, "h5", "h5io.cpp"), "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib])
    os.replace(tmp, H5_LIB)
    if verbose:
        print("built", H5_LIB)
    return H5_LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "engine_test.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "engine_test")


def build_cpp_test(verbose=False):
    """the C++ program that exercises include/qcmaquis_amd.hpp against the library (run by
    tests/test_cpp_engine.py on the GPU)"""
    tmp = CPP_TEST_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", tmp, CPP_TEST_SRC,
          "-L" + PKG, "-lqcmaquis_amd", "-Wl,-rpath,$ORIGIN/../../qcmaquis_amd"])
    os.replace(tmp, CPP_TEST_BIN)
    if verbose:
        print("built", CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    build_library(verbose=True)
    build_h5(verbose=True)
    build_cpp_test(verbose=True)
