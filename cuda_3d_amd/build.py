"""Build libgca.so (HIP, gfx950) in-tree: cuda_3d_amd/libgca.so.

    python cuda_3d_amd/build.py          # compile every source, link
    python cuda_3d_amd/build.py --check  # report whether the library is present
(run it as a script: importing the package loads the library it is about to replace)

Flags: -ffp-contract=off keeps every floating-point operation a single IEEE rounding (the
canonical discrete rules and the bit-exact FP64 paths depend on it); packed FP32 ops are
disabled.  Sources compile in parallel; each object is rebuilt every call (the build is the
"does it build" check).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgca.so")
SOURCES = ["api.hip", "frontend.hip", "voxel.hip", "sort.hip", "grid.hip", "normals.hip", "icp.hip", "pipeline.hip", "cluster.hip",
           "noalloc.hip"]
ARCH = os.environ.get("GCA_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        sys.stderr.write("\n".join(warn[:20]) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(LIB if os.path.exists(LIB) else "missing")
    else:
        build()
