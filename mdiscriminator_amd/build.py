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
    "arrays.hip",
    "segmented.hip",
    "serve.cpp",
] + [f"seg_cfg_b{b}.hip" for b in range(7)]  # segmented candidates, one unit per class-bit count

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
    "-pthread",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build_library(verbose=False):
    """Compiles every source to an object in parallel, then links the library."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs = []
    objs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        # serve.cpp is host code; -x hip gives it the same HIP runtime headers.
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmd = [_hipcc()] + compile_flags + ["-c"] + lang + [src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building libmdiscr.so")
    tmp = LIB + ".tmp"
    cmd = [_hipcc()] + HIPCC_FLAGS + ["-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libmdiscr.so failed")
    os.replace(tmp, LIB)
    return LIB


PROBE_LIB = os.path.join(PKG, "libclkprobe.so")


def build_probe(verbose=False):
    """The shader-clock probe bench.py --first-steps samples between steps."""
    cmd = [_hipcc()] + HIPCC_FLAGS + ["-o", PROBE_LIB, os.path.join(CSRC, "clock_probe.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building libclkprobe.so failed")


HOST_PROGRAMS = {"mdiscr_segbench": "mdiscr_segbench.c", "mdiscr_servebench": "mdiscr_servebench.c"}
BIN = os.path.join(PKG, "bin")


def build_host_programs(verbose=False):
    """C host drivers of the C ABI (linked against libmdiscr.so)."""
    os.makedirs(BIN, exist_ok=True)
    for exe, src in HOST_PROGRAMS.items():
        cmd = ["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-D__HIP_PLATFORM_AMD__",
               "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
               "-o", os.path.join(BIN, exe), os.path.join(PKG, "host", src),
               "-L", PKG, "-lmdiscr", "-L", "/opt/rocm/lib", "-lamdhip64", "-lm", "-pthread",
               "-Wl,-rpath,$ORIGIN/..:/opt/rocm/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building {exe} failed")


def build_test_programs(verbose=False):
    """tests/c/mdiscr_ref_driver: the C++ mirror of the reference templates,
    checked against the oracle (built after the oracle)."""
    out = os.path.join(ROOT, "tests", "c", "mdiscr_ref_driver")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host"),
           "-I", os.path.join(ROOT, "oracle"), "-I", "/opt/rocm/include",
           "-o", out, os.path.join(ROOT, "tests", "c", "mdiscr_ref_driver.cpp"),
           "-L", PKG, "-lmdiscr", "-L", os.path.join(ROOT, "oracle"), "-loracle",
           "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../mdiscriminator_amd:$ORIGIN/../../oracle:/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building tests/c/mdiscr_ref_driver failed")


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
    build_probe(verbose)
    build_host_programs(verbose)
    build_oracle(verbose)
    build_test_programs(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
