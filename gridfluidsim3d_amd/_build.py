"""Builds libgfsx.so (the HIP library behind include/gfsx.h) in-tree for gfx950.

hipcc cross-compiles for gfx950 without a GPU.  Packed-FP32 instructions are
disabled for every HIP translation unit (-Xclang -target-feature -Xclang
-packed-fp32-ops); hipcc's host pass warns that it does not know the feature,
which is harmless.
"""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libgfsx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["gfsx_kernels.hip", "gfsx_adjoint.hip", "gfsx_epilogue.hip", "gfsx_api.hip"]

HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    """Compile every HIP source of the library into lib/libgfsx.so."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ([HIPCC] + HIP_FLAGS + ["-shared", "-Wl,-soname," + os.path.basename(LIB),
                                  "-I", os.path.join(ROOT, "include"),
                                  "-I", CSRC, "-o", tmp] +
           [os.path.join(CSRC, s) for s in SOURCES])
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
