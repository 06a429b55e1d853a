"""Build libcvt_hip.so (gfx950) in-tree with hipcc.

Usage: python -m gpu_cvt_amd.build [--jobs N]
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
# CVT_STAGE_CLOCKS=1: the profiling build with per-stage shader clocks
# (-DCVT_STAGE_CLOCKS) into its own library, beside the product one
STAGE_CLOCKS = os.environ.get("CVT_STAGE_CLOCKS", "0") == "1"
LIB = os.path.join(HERE, "libcvt_hip_stageclk.so" if STAGE_CLOCKS else "libcvt_hip.so")
CLI = os.path.join(HERE, "gpu_cvt")
SOURCES = ["cvt_kernels.hip", "cvt_bvh.hip", "cvt_api.hip", "cvt_server.hip"]
# cvt_lloyd.hip is compiled once per neighbour count K (-DCVT_K=k) and kproj (-DCVT_KP=p)
PER_K_SOURCE = "cvt_lloyd.hip"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CVT_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep hipcc from emitting packed FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result", "-pthread",
]


def k_values() -> list[int]:
    """The K list of CVT_FOR_EACH_K in cvt_kernels.h (the run-time dispatch)."""
    with open(os.path.join(CSRC, "cvt_kernels.h")) as f:
        for line in f:
            if line.startswith("#define CVT_FOR_EACH_K(X)"):
                return [int(t) for t in re.findall(r"X\((\d+)\)", line)]
    raise RuntimeError("CVT_FOR_EACH_K not found in cvt_kernels.h")


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_obj_stageclk" if STAGE_CLOCKS else "_obj")
    os.makedirs(objdir, exist_ok=True)
    units = [(src, src.replace(".hip", ".o"), []) for src in SOURCES]
    units += [(PER_K_SOURCE, f"cvt_lloyd_k{k}.o", [f"-DCVT_K={k}"]) for k in k_values()]
    # the FP64 pass with kproj = 10 lists, one object per K
    units += [(PER_K_SOURCE, f"cvt_lloyd_k{k}_p10.o", [f"-DCVT_K={k}", "-DCVT_KP=10"]) for k in k_values()]

    def compile_one(unit) -> str:
        src, oname, defs = unit
        obj = os.path.join(objdir, oname)
        extra = ["-DCVT_STAGE_CLOCKS"] if STAGE_CLOCKS else []
        cmd = [HIPCC, *FLAGS, *extra, *defs, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    jobs = int(os.environ.get("MAX_JOBS", "0")) or min(16, os.cpu_count() or 1)
    # the largest units first, so the pool drains evenly
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sorted(units, key=lambda u: u[0] != PER_K_SOURCE)))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if STAGE_CLOCKS:
        return LIB
    # the `gpu_cvt` command line (host C++ over the C ABI)
    cli_tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "cvt_cli.cpp"),
           "-I", os.path.join(ROOT, "include"), "-L", HERE, "-lcvt_hip", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", cli_tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(cli_tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
