"""Compile libqmctoolscl_amd.so for gfx950 in place (next to this file).

    python -m qmctoolscl_amd.build [-j N]

Each source is compiled by hipcc to an object (in parallel), then linked into
one shared library exporting the C ABI declared in include/qmctoolscl_amd.h.
Floating-point contraction is disabled library-wide: the kernels reproduce
the reference C backend's unfused double arithmetic bit for bit.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqmctoolscl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["capi.cpp", "lattice.hip", "dnb2.hip", "gdn.hip", "transforms.hip", "nus.hip", "mfma_fwht.hip", "mfma_fft.hip", "batch.hip", "halton.hip", "kernel_methods.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-ffp-contract=off",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    out = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    return out


def build(jobs=8, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    try:
        build(args.j)
    except RuntimeError as e:
        sys.exit(str(e))
