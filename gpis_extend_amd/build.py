"""Builds the in-tree HIP library libgpis_hip.so for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(PKG, "csrc", "gpis_hip.hip")
OUT = os.path.join(PKG, "libgpis_hip.so")

# Packed FP32 forms are disabled for every HIP object of this project.  The two-address
# v_fmac_f64 form is disabled too: with it the compiler copies each loop-invariant FP64
# coefficient into the accumulator before every fused multiply-add (a v_mov_b64 per
# Horner step of the splat exponential); v_fma_f64 takes the coefficient as a third source.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-pass-failed",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


HOST_DIR = os.path.join(PKG, "host")
SELFTEST = os.path.join(HOST_DIR, "backend_selftest")
POOLTEST = os.path.join(HOST_DIR, "path_pool_selftest")


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    # the C++ batch backend (host code over the C ABI) and its self-test driver
    cmd = [hipcc(), "-O2", "-std=c++17", "-Wno-unused-result",
           os.path.join(HOST_DIR, "gpis_backend.cpp"), os.path.join(HOST_DIR, "backend_selftest.cpp"),
           "-L" + PKG, "-lgpis_hip", "-Wl,-rpath,$ORIGIN/..", "-lpthread", "-o", SELFTEST + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(SELFTEST + ".tmp", SELFTEST)
    # the wavefront path pool over the backend and its self-test driver (no contraction:
    # the host's hit-time phase rounds as the device's)
    cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unused-result",
           os.path.join(HOST_DIR, "gpis_backend.cpp"), os.path.join(HOST_DIR, "path_pool.cpp"),
           os.path.join(HOST_DIR, "path_pool_selftest.cpp"),
           "-L" + PKG, "-lgpis_hip", "-Wl,-rpath,$ORIGIN/..", "-lpthread", "-o", POOLTEST + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(POOLTEST + ".tmp", POOLTEST)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
