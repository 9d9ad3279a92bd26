"""Builds the native libraries of proresraw_rs_amd in-tree.

- ``_lib/libproresraw_amd.so``: the HIP kernels (decode, develop, learned
  demosaic, DNG tile coding, crop and orientation), the C-ABI host layer and the movie
  sample-table reader,
  compiled for gfx950 with hipcc;
- ``_lib/libprr_synth.so``: the synthetic frame generator (plain C).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Floating point exactly as the reference crate: no contraction into
    # fused multiply-adds beyond the explicit fmaf calls, IEEE denormals,
    # correctly rounded division and square root.
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wshadow",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libproresraw_amd.so")
    srcs = [os.path.join(CSRC, n) for n in
            ("prr_kernels.hip", "prr_net.hip", "prr_dng.hip", "prr_orient.hip", "prr_host.cpp",
             "prr_parse.cpp", "prr_net_host.cpp", "prr_mov.cpp")]
    tmp = out + ".tmp"
    _run(["/opt/rocm/bin/hipcc", *HIPCC_FLAGS, *srcs, "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


def build_synth(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libprr_synth.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-pthread", "-Wall",
          os.path.join(CSRC, "synth.c"), "-o", tmp, "-lm"], verbose)
    os.replace(tmp, out)
    return out


def build(verbose=True):
    build_synth(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
