"""Where the FP32 Lloyd kernels spend their time, by stage, on one workload.

Needs the profiling build of the library, which adds per-stage shader-clock
sums to the product kernels (compiled with -DCVT_STAGE_CLOCKS; the product
build has none of it):

    CVT_STAGE_CLOCKS=1 python -m gpu_cvt_amd.build
    CVT_LIB=gpu_cvt_amd/libcvt_hip_stageclk.so python -m gpu_cvt_amd.stage_split [--workload W] [--warmup N] [--iters M]

Each iteration runs kernel by kernel with HIP events around each kernel
(cvt_job_profile_stages); every lane adds its cycles per stage.  A kernel's
event time is split over its stages in proportion to their cycle sums, so the
per-stage microseconds add up to the kernel's time (cycles include the time a
wave waits while other waves of its SIMD issue).  Prints a markdown table.
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys

import numpy as np

STAGES = [  # (stage, kernel) in the order of the clock sums
    ("site KNN: seeds, collect, selection, order proof", "cell"),
    ("KNN row write and dual points", "cell"),
    ("dual hull walk and cell moments", "cell"),
    ("vertex list: pool cache or grid search", "project"),
    ("faces: nearest point, new site, freeze step, its nearest vertex", "project"),
    ("nearest vertex and tangent frame (frame kernel: set-up only)", "frame"),
]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="dsphere_2m_1m")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--iters", type=int, default=5)
    args = ap.parse_args(argv)
    from gpu_cvt_amd import cvt, meshgen as mg

    L = cvt.lib()
    if not hasattr(L, "cvt_stage_clocks"):
        print(f"{cvt.LIB_PATH} is not the stage-clock build (CVT_STAGE_CLOCKS=1 python -m gpu_cvt_amd.build, "
              "then CVT_LIB=gpu_cvt_amd/libcvt_hip_stageclk.so)", file=sys.stderr)
        return 2
    L.cvt_stage_clocks.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong)]
    V, F, S, K, _iters, _fz = mg.workload(args.workload)
    print(f"[stage_split] {args.workload}: nS={len(S)} nF={len(F)} K={K}", file=sys.stderr, flush=True)
    clk = np.zeros(len(STAGES))
    counts = np.zeros(4)
    ms = {"cell": 0.0, "project": 0.0, "frame": 0.0, "grid": 0.0, "precise": 0.0, "iter": 0.0}
    with cvt.Job(V, F, S, K=K) as job:
        far = [r["far_walks"] for r in job.step(args.warmup, stats=True)]
        for _ in range(args.iters):
            st = job.profile_stages(1)
            out = (C.c_ulonglong * (len(STAGES) + len(counts)))()
            if L.cvt_stage_clocks(job.handle, out) != 0:
                raise RuntimeError(L.cvt_last_error().decode())
            clk += np.array(out[:len(STAGES)], dtype=np.float64)
            counts += np.array(out[len(STAGES):], dtype=np.float64)
            for k in ms:
                ms[k] += st[k] / args.iters
    print(f"# Stage split, {args.workload} (iterations {args.warmup + 1}-{args.warmup + args.iters}, "
          f"mean of {args.iters})\n")
    print("| kernel | stage | share of kernel cycles | us per iteration |")
    print("|---|---|---|---|")
    for kern in ("cell", "project"):
        idx = [i for i, (_, k) in enumerate(STAGES) if k == kern]
        tot = clk[idx].sum()
        for i in idx:
            share = clk[i] / tot if tot > 0 else 0.0
            print(f"| kb_{kern} ({1e3 * ms[kern]:.1f} us) | {STAGES[i][0]} | {100 * share:.1f} % | "
                  f"{1e3 * ms[kern] * share:.1f} |")
    print(f"\nOther stages per iteration: site-grid rebuild {1e3 * ms['grid']:.1f} us, FP64 pass "
          f"{1e3 * ms['precise']:.1f} us; whole iteration (kernel by kernel, with events) {1e3 * ms['iter']:.1f} us.")
    n = counts / args.iters
    print(f"\nPer iteration: projection kernel vertex lists {n[0]:.0f}, of which the pool cache could not prove "
          f"{n[1]:.0f} ({100 * n[1] / max(n[0], 1):.1f} %: a grid search); new-site nearest-vertex searches "
          f"(neither the projection's list nor the pool proved it) {n[2]:.0f}; cell kernel site searches the LDS collect did not prove {n[3]:.0f}.")
    print(f"\nFP32 vertex searches finished by the out-of-line LBVH walk (the FP32 kernels' only code with a "
          f"stack frame), iterations 1-{args.warmup}: {far}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
