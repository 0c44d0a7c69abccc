"""Build libsht.so (gfx950 kernels + host plan code + C ABI) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsht.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in
           ("sht_legendre.hip", "sht_fft.hip", "sht_dft.hip", "sht_spectral.hip", "sht_mlp.hip", "sht_static.hip", "sht_plan.cpp")]

ROCM_LIB = "/opt/rocm/lib"  # roctx ranges (rocprofiler-sdk-roctx)
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay disabled in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build_library(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, "-I", os.path.join(REPO, "include"), "-x", "hip", *SOURCES,
           "-L", ROCM_LIB, "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + ROCM_LIB, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
