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

SOURCES = ["gfsx_kernels.hip", "gfsx_brick.hip", "gfsx_adjoint.hip", "gfsx_epilogue.hip", "gfsx_remove.hip", "gfsx_api.hip"]

HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    """Compile every HIP source of the library (in parallel, one object per
    source) and link them into lib/libgfsx.so."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = ([HIPCC] + HIP_FLAGS + ["-c", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                                      "-o", obj, os.path.join(CSRC, src)])
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((obj, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True)))
    errors = []
    for obj, cmd, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            errors.append(" ".join(cmd) + "\n" + out + err)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = ([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC",
            "-Wl,-soname," + os.path.basename(LIB), "-o", tmp] + [obj for obj, _, _ in jobs])
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
