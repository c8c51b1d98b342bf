"""Build libcfs_plume.so (HIP kernels for gfx950 + host C ABI) in-tree.

    python -m cfs_plume_amd.build

HIP sources are compiled by hipcc for gfx950 with packed-FP32 instructions
disabled; the host encoder is compiled by g++ without FMA contraction so that
its FP64 arithmetic is reproducible.  Objects go to cfs_plume_amd/_build/, the
library to cfs_plume_amd/libcfs_plume.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcfs_plume.so")
ARCH = "gfx950"

HIP_SOURCES = ["step_kernels.hip", "decode_kernels.hip", "closure_kernels.hip", "noalloc.hip", "cubature.hip", "plume.hip", "fluid.hip", "linalg.hip", "obstacle.hip"]
HIP_HOST_SOURCES = ["basis.cpp", "batch.cpp", "closure.cpp"]
CXX_SOURCES = ["encoder.cpp", "nnls.cpp"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=fast",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP extension cannot be built")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([hipcc, "-O2", "-std=c++17", "-fPIC", "-Wall", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
                     "-I/opt/rocm/include", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    procs = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)) for cmd in jobs]
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"FAILED: {' '.join(cmd)}\n{out}\n")
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("building libcfs_plume.so failed")
    tmp = LIB + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(link, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
