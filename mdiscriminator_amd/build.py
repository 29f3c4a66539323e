"""Builds the gfx950 library (libmdiscr.so, in-tree) and the oracle.

    python -m mdiscriminator_amd.build          # library + oracle
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmdiscr.so")

SOURCES = [
    "mdiscr_api.hip",
    "soa.hip",
    "segmented.hip",
]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # Packed FP32 VALU forms are kept out of every code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [_hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmdiscr.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")],
                       capture_output=True, text=True)
    if verbose:
        sys.stdout.write(r.stdout)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")


def build_all(verbose=False):
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
