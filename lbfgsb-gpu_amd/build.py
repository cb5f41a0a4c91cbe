"""Build liblbfgsb_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

    python lbfgsb-gpu_amd/build.py [--jobs N]

Every source is compiled on every call; objects go to lbfgsb-gpu_amd/_build
and the library to lbfgsb-gpu_amd/lib/liblbfgsb_amd.so.  The example of the
C++ drop-in API (examples/dsscfg.hip, over include/culbfgsb.h) is linked
against it as lbfgsb-gpu_amd/lib/dsscfg_example.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "liblbfgsb_amd.so")
EXAMPLE_SRC = os.path.join(HERE, "examples", "dsscfg.hip")
EXAMPLE = os.path.join(LIBDIR, "dsscfg_example")

SOURCES = ["batch.hip", "bulk.hip", "small.hip", "gcp.hip", "freev.hip", "objectives.hip",
           "generate.hip", "noalloc.hip", "engine.cpp", "capi.cpp", "batch_host.cpp", "serve.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LBFGSB_AMD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + CSRC, "-I" + os.path.join(REPO, "include"),
]
# device code: gfx950 only; packed FP32 instructions are kept out of the code
# object (they are not used by hand anywhere either)
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON
    cmd += DEVICE if src.endswith(".hip") else ["-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [HIPCC] + COMMON + DEVICE + [EXAMPLE_SRC, "-x", "none", "-L" + LIBDIR,
                                       "-llbfgsb_amd", "-Wl,-rpath,$ORIGIN", "-o", EXAMPLE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed:\n{r.stderr}")
    if verbose:
        print("built", LIB, "and", EXAMPLE)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
