"""Builds the HIP library for gfx950 in-tree:

    python -m pamhd_amd.build     ->  pamhd_amd/lib/libpamhd_mi355x.so
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libpamhd_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the FP64 path must round every product and sum
# separately, as the CPU reference does.  Packed FP32 instructions are
# disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build_library(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in
            ("pamhd_mi355x.hip", "poisson.hip", "swbox.hip", "adjoint.hip", "plain1d.hip",
             "ensemble.cpp", "driver.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                 "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


def build(verbose=False):
    return build_library(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
