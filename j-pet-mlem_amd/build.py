"""Builds libpetstrip_hip.so (HIP kernels + C ABI) for gfx950, in-tree, and
libpetstrip_probe.so (the atomic-throughput probes bench.py reports).
Host-only sources (the response-file reader) are compiled by the host C++
compiler into the same library.

Every HIP source is compiled with -ffp-contract=off: the few fused
multiply-adds the strip geometry needs are written out explicitly, so the
discrete pixel decisions follow the reference CPU build exactly.  Packed
FP32 instructions are disabled (-packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpetstrip_hip.so")
PROBE_LIB = os.path.join(HERE, "libpetstrip_probe.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PETSTRIP_ARCH", "gfx950")

SOURCES = ["strip_kernels.hip", "strip_bins.hip", "strip_fused.hip", "petstrip_capi.hip",
           "phantom.hip", "hybrid.hip", "noalloc.hip"]
PROBE_SOURCES = ["atomic_probe.hip"]
HOST_SOURCES = ["text_responses.cpp", "system_matrix.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall",
              "-I", os.path.join(ROOT, "include"), "-I", CSRC]

FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    """Compiles every source (always) and links the two shared libraries."""
    os.makedirs(BUILD, exist_ok=True)
    cmds = []
    objs = {}
    for src in SOURCES + PROBE_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs[src] = obj
        cmds.append([HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        objs[src] = obj
        cmds.append([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    for lib, srcs in ((LIB, SOURCES + HOST_SOURCES), (PROBE_LIB, PROBE_SOURCES)):
        tmp = lib + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
              *[objs[s] for s in srcs], "-lpthread"])
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
