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
         ["-L/opt/rocm/lib", "-lrocblas"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
