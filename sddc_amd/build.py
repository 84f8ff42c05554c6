"""Build the device library (hipcc, gfx950) in-tree.

    python -m sddc_amd.build          # builds sddc_amd/lib/libsd2d.so

The library is plain HIP C++ behind the C ABI of include/sd2d.h; nothing
here depends on torch.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsd2d.so")
SRC = os.path.join(HERE, "csrc", "sd2d.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-function",
    # no packed-FP32 instructions anywhere (they are not used by FP64 code
    # either, but the feature stays off for every HIP build of this repo)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_library(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIPFLAGS, "-I", os.path.join(REPO, "include"), SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libsd2d.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
    return os.path.join(REPO, "oracle", "liboracle_sd2d.so")


def build_reference() -> bool:
    """Out-of-tree build of the reference Fortran (only where its sources
    are present), and of the reference program with its compresid replaced
    by this library (needs the library built first).  Returns False when
    there is nothing to build."""
    if not os.path.isdir("/root/reference"):
        return False
    subprocess.run(["make", "-s", "-f", os.path.join(REPO, "oracle", "ref.mk"), "all", "dropin"],
                   check=True, cwd=REPO, stdout=subprocess.DEVNULL)
    return True


if __name__ == "__main__":
    print(build_library(verbose=True))
