"""Builds libsfft_amd.so in-tree for gfx950 with hipcc.

Host sources (*.cpp) are compiled as plain C++; kernels (*.hip) as HIP for
gfx950.  Everything is compiled with -ffp-contract=off: the FP64 form
reproduces the reference's separately rounded products and sums.  Packed
FP32 instructions are disabled for the device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsfft_amd.so")

HOST_SRCS = ["host_params.cpp", "host_filters.cpp", "host_plan.cpp", "api.cpp", "server.cpp",
             "tune.cpp", "host_signal.cpp"]
DEVICE_SRCS = ["kernels_v12.hip", "kernels_v3.hip", "kernels_fft.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-I" + os.path.join(HERE, "..", "include")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    hipcc = _hipcc()
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    # hipcc compiles .cpp as HIP too; every source targets gfx950 only.
    cmd = [hipcc] + COMMON + ["-x", "hip"] + DEVICE
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = HOST_SRCS + DEVICE_SRCS
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
