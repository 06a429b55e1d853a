"""Benchmark: Lloyd site-iterations/s of the HIP surface-CVT path on one MI355X.

Workload (default): displaced cube-sphere with 1,997,568 triangles (998,786
vertices), 1,000,000 area-uniform sites, K = 24, freezing off, FP32 -- the
large-job shape the headline metric is quoted on.  A "step" is one Lloyd
iteration over all sites (site-grid rebuild, then the frame, cell and
projection kernels), launched as one HIP graph; inputs are resident in HBM
before the timed region.

Prints one JSON line (rank 0).  Also measures, in the same run:
  * roofline: algorithmic bytes of one Lloyd iteration divided by the mean
    duration of the FP32 Lloyd pass -- the two kernels that do all of the
    per-site work, kb_cell + kb_project, launched back to back (the frame
    vertex of each new site is found by the projection kernel) --
    timed with HIP events on the job's stream; beside it valu_frac, the pass's
    VALU issue floor (from the committed SQ-counter profile) over the same time;
  * cpu_baseline: an FP32 C++ Lloyd step and the FP64 CPU oracle's step on this
    host's allowed CPUs, on a bounded sample (a sub-sampled site set on the
    same mesh, or the whole job over as many iterations as fill about half of
    --cpu-budget seconds);
  * unless --no-extra: freezing on, K = 8, the batched form (256 jobs) and the
    served form (a Poisson request trace) on the same GPU.
Other modes: batch, serve, check (see --mode).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Lloyd site-iters/s at 1M sites (% HBM roofline); p50/p99 remesh latency"


def algorithmic_bytes(nS: int, nV: int, nF: int, K: int) -> int:
    """Compulsory bytes of one Lloyd iteration with freezing off: each site read
    (12 B) and written (12 B), its K-list written (4K B), and the mesh read once
    (positions 12 nV, normals 12 nV, faces 12 nF, vertex->face CSR 4(nV+1) + 12 nF)."""
    return nS * (24 + 4 * K) + 28 * nV + 24 * nF + 4


def traffic_from_profile(workload: str):
    """Bytes past the L2 per FP32 Lloyd pass (kb_cell + kb_project),
    as recorded by the committed rocprofv3 summary of this workload (reads from
    the per-request-size L2 read counters, writes from WRITE_SIZE); (None,
    None) when no summary exists for the workload.  The counters cannot be read
    from inside this process, so the figure comes from the profile run of the
    same command."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"lloyd_pass_{workload}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1].strip()), os.path.relpath(path, os.path.dirname(path) + "/..")
    return None, None


def valu_issue_from_profile(workload: str):
    """The VALU issue floor of the FP32 Lloyd pass in microseconds per launch
    (VALU instructions per wave x 4 cycles x waves per SIMD at 2.4 GHz, summed
    over kb_cell and kb_project), as recorded by the committed
    rocprofv3 SQ-counter summary of this workload; None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"lloyd_pass_{workload}_sq_counters.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("valu_issue_us_per_launch:"):
                return float(line.split(":")[1].strip())
    return None


def cpu_threads() -> tuple[int, int]:
    """(threads used, hardware threads): every CPU this process may run on,
    capped at the machine's CPU share for this process when the environment
    states one (OMP_NUM_THREADS)."""
    hw = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else hw
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(allowed, share) if share > 0 else allowed), hw


def _time_cpu(step, n_total: int, budget_s: float, sub_sizes):
    """Times step(n, iters) -> seconds on growing site samples until one run
    fills about half the budget; returns (sites, iterations, seconds)."""
    n, it, dt = sub_sizes[0], 1, step(sub_sizes[0], 1)
    for n2 in sub_sizes[1:]:
        if dt > 0.25 * budget_s:
            break
        n, dt = n2, step(n2, 1)
    if n == n_total and dt < 0.5 * budget_s:
        it = max(1, int(0.5 * budget_s / dt))
        if it > 1:
            dt = step(n, it)
    return n, it, dt


def cpu_baseline(V, F, S, K: int, budget_s: float = 20.0):
    """The CPU Lloyd step on this host, on bounded samples of the same job:
    the FP32 speed variant (oracle/cvt_cpu32.cpp) and the FP64 oracle, each on
    every CPU this process may use (pinned), about budget_s / 2 each."""
    from oracle import oracle as orc

    orc.build()
    threads, hw = cpu_threads()
    orc.set_threads(threads)
    mesh = orc.Mesh(V, F)
    R = float(np.float32(mesh.bbox_max_edge))
    rng = np.random.Generator(np.random.PCG64(123))
    sizes = sorted({min(len(S), x) for x in (20_000, 100_000, 300_000, len(S))})
    subs = {n: S[np.sort(rng.choice(len(S), n, replace=False))] if n < len(S) else S for n in sizes}

    def fp32(n, iters):
        t0 = time.perf_counter()
        orc.cpu32_lloyd(mesh, subs[n], K, R, iters, threads)
        return time.perf_counter() - t0

    def fp64(n, iters):
        st = orc.LloydState(subs[n], K)
        t0 = time.perf_counter()
        for _ in range(iters):
            orc.lloyd_step(mesh, st, K=K, R=R)
        return time.perf_counter() - t0

    n32, it32, dt32 = _time_cpu(fp32, len(S), 0.5 * budget_s, sizes)
    n64, it64, dt64 = _time_cpu(fp64, len(S), 0.5 * budget_s, sizes)
    v32, v64 = n32 * it32 / dt32, n64 * it64 / dt64
    model = orc.cpu_model()
    return {"value": v32, "unit": "site-iters/s", "cores": threads, "kind": "port",
            "sample": f"FP32 C++ Lloyd step: {it32} iteration(s) over {n32} of the {len(S)} sites ({dt32:.1f} s); "
                      f"FP64 oracle step: {it64} iteration(s) over {n64} sites ({dt64:.1f} s); same mesh, "
                      f"{threads} pinned threads",
            "threads": threads, "hardware_threads": hw, "model": model, "fp32_value": v32, "fp64_value": v64}


def run_batch(jobs: int, iters: int, warmup: int) -> dict:
    """Mixed-size jobs (log-uniform triangles 5k-400k, sites 1k-100k capped at
    triangles/2, K = 24, freezing on): site-iters/s of the batched form (all
    jobs in one graph launch per iteration) against the plain path (each job's
    own graph, one job at a time) on the same inputs.  Both forms compute
    bit-identical results (tests/test_gpu_batch.py), so the final Hausdorff is
    equal by construction."""
    from gpu_cvt_amd import cvt, meshgen as mg

    t_gen = time.perf_counter()
    work = mg.batch_workload(jobs, seed=7)
    t_gen = time.perf_counter() - t_gen
    total_sites = sum(len(w[2]) for w in work)
    kw = dict(K=24, freeze_mode=1, freeze_need=5, freeze_disp_rel=1e-5)
    print(f"[bench] batch: {len(work)} jobs, {total_sites} sites, "
          f"{sum(len(w[1]) for w in work)} triangles, gen {t_gen:.1f}s", file=sys.stderr, flush=True)
    # plain path: one job at a time, each job's own graph
    plain_ms = 0.0
    for k, (V, F, S, _tag) in enumerate(work):
        with cvt.Job(V, F, S, **kw) as job:
            job.step(warmup)
            plain_ms += job.step_timed(iters)
        if k % 32 == 31:
            print(f"[bench] plain: {k + 1} jobs done", file=sys.stderr, flush=True)
    plain_rate = total_sites * iters / (plain_ms / 1e3)
    # batched: all jobs resident, one graph launch per iteration
    jobs = [cvt.Job(V, F, S, **kw) for V, F, S, _tag in work]
    with cvt.Batch(jobs) as b:
        b.step(warmup)
        batch_ms = b.step_timed(iters)
    frozen = sum(int(j.freeze_state()[0].sum()) for j in jobs)
    for j in jobs:
        j.close()
    batch_rate = total_sites * iters / (batch_ms / 1e3)
    print(f"[bench] plain {plain_ms:.1f} ms, batched {batch_ms:.1f} ms for {iters} iterations; "
          f"frozen at end {frozen}/{total_sites}", file=sys.stderr, flush=True)
    return {"jobs": len(work), "sites": total_sites, "triangles": int(sum(len(w[1]) for w in work)), "K": 24,
            "freezing": True, "iterations": iters, "warmup": warmup, "batched_site_iters_per_s": batch_rate,
            "plain_site_iters_per_s": plain_rate, "batched_ms_per_iteration": batch_ms / iters,
            "batched_over_plain": batch_rate / plain_rate, "frozen_at_end": frozen}


def bench_batch(args) -> int:
    """Batched form (the cfg4 shape), printed as its own JSON line."""
    r = run_batch(args.jobs, args.steps, args.warmup)
    out = {
        "metric": "batched Lloyd site-iters/s, 256 mixed jobs (vs one job at a time)",
        "value": r["batched_site_iters_per_s"], "unit": "site-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["batched_ms_per_iteration"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded icospheres, tori, displaced spheres)",
        "config": {"workload": "batch_mixed", "jobs": r["jobs"], "sites": r["sites"], "K": 24, "freezing": True},
        "plain_value": r["plain_site_iters_per_s"], "batched_over_plain": r["batched_over_plain"],
    }
    print(json.dumps(out), flush=True)
    return 0


def run_serve(arrivals: int, rate: float, ingest_threads: int = 4, serve_log: str = "") -> dict:
    """Serving form: a seeded Poisson request trace (arrivals at `rate` jobs/s,
    batched-form sizes plus 5% large 1M-site jobs) submitted in real time to
    the continuous-batching loop (cvt_server_*) under a 200 GB admission
    budget; each job runs to a one-sided Hausdorff distance <= 1e-3 x bbox
    diagonal, checked every 10 iterations, or 100 iterations.  The same trace
    is then run first-in-first-out, one job at a time, on the plain path
    (cvt_job_step + cvt_job_hausdorff_within) with the same checks.  Both paths
    give every job bit-identical iterates, so the final Hausdorff values agree
    exactly.  Returns p50/p99 time-to-convergence (arrival -> result) over the
    converged jobs and over all jobs, and aggregate site-iters/s, of both."""
    from gpu_cvt_amd import cvt, meshgen as mg

    t_gen = time.perf_counter()
    trace = mg.serve_trace(arrivals, rate, seed=8)
    t_gen = time.perf_counter() - t_gen
    kw = dict(K=24, freeze_mode=1, freeze_need=5, freeze_disp_rel=1e-5)
    tol, max_iters = 1e-3, 100
    n_large = sum(1 for x in trace if x[4].startswith("large"))
    print(f"[bench] serve: {len(trace)} requests over {trace[-1][0]:.1f} s ({n_large} large), "
          f"{sum(len(x[3]) for x in trace)} sites, gen {t_gen:.1f}s", file=sys.stderr, flush=True)

    # warm both code paths (module load, first graph instantiation) on a small job
    V0, F0, S0 = trace[0][1], trace[0][2], trace[0][3]
    with cvt.Job(V0, F0, S0, **kw) as j0:
        j0.step(10)
        j0.hausdorff()

    # served: continuous batching
    with cvt.Server(hbm_budget_bytes=200 << 30, epoch_iters=10, min_check_iters=10, check_disp_rel=1e30,
                    check_frozen_frac=0.0, ingest_threads=ingest_threads) as srv:
        t0 = time.perf_counter()
        tickets = []
        for k, (ta, V, F, S, _tag) in enumerate(trace):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(V, F, S, tol_rel=tol, max_iters=max_iters, **kw))
            if k % 500 == 499:
                print(f"[bench] serve: {k + 1} submitted, info {srv.info()}", file=sys.stderr, flush=True)
        srv.wait_idle(3000)
        info = srv.info()
        res = [srv.poll(t) for t in tickets]
    if serve_log:
        with open(serve_log, "w") as f:
            json.dump([{k: r[k] for k in ("nS", "iterations", "state", "t_arrive", "t_admit", "t_done")} for r in res], f)
    lat_s = np.array([r["t_done"] - r["t_arrive"] for r in res])
    span_s = max(r["t_done"] for r in res) - min(r["t_arrive"] for r in res)
    si_s = float(sum(r["nS"] * r["iterations"] for r in res))
    h_s = np.array([r["hausdorff"] for r in res])
    conv = np.array([r["state"] == "converged" for r in res])
    print(f"[bench] served: span {span_s:.2f} s, p50 {np.percentile(lat_s, 50):.3f} s, "
          f"p99 {np.percentile(lat_s, 99):.3f} s, converged {int(conv.sum())}/{len(res)}, info {info}",
          file=sys.stderr, flush=True)

    # plain path, FIFO, one job at a time
    t0 = time.perf_counter()
    lat_f, h_f, si_f, it_f = [], [], 0.0, []
    for k, (ta, V, F, S, _tag) in enumerate(trace):
        dt = ta - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        with cvt.Job(V, F, S, **kw) as job:
            diag = job.info()["bbox_diag"]
            it, h = 0, -1.0
            while True:
                job.step(10)
                it += 10
                ok, h = job.hausdorff_within(tol * diag)
                if ok:
                    break
                if it >= max_iters:
                    h = job.hausdorff()[0]
                    break
        lat_f.append(time.perf_counter() - t0 - ta)
        h_f.append(h)
        it_f.append(it)
        si_f += len(S) * it
        if k % 500 == 499:
            print(f"[bench] fifo: {k + 1} done", file=sys.stderr, flush=True)
    span_f = time.perf_counter() - t0 - trace[0][0]
    lat_f = np.array(lat_f)
    same_iters = all(r["iterations"] == i for r, i in zip(res, it_f))
    h_diff = float(np.max(np.abs(h_s - np.array(h_f))))
    print(f"[bench] fifo: span {span_f:.2f} s, p50 {np.percentile(lat_f, 50):.3f} s, "
          f"p99 {np.percentile(lat_f, 99):.3f} s; same iteration counts {same_iters}, max |h diff| {h_diff:.3e}",
          file=sys.stderr, flush=True)

    def pct(x, q):
        return float(np.percentile(x, q)) if len(x) else None

    return {
        "requests": len(trace), "rate_per_s": rate, "large": n_large, "converged": int(conv.sum()),
        "reached_max_iters": int(len(res) - conv.sum()), "K": 24, "freezing": True, "tol_hausdorff_rel": tol,
        "check_every": 10, "max_iters": max_iters, "hbm_budget_gb": 200,
        "site_iters_per_s": si_s / span_s, "span_s": span_s,
        "p50_s": pct(lat_s, 50), "p99_s": pct(lat_s, 99),
        "p50_converged_s": pct(lat_s[conv], 50), "p99_converged_s": pct(lat_s[conv], 99),
        "fifo": {"site_iters_per_s": si_f / span_f, "span_s": span_f, "p50_s": pct(lat_f, 50),
                 "p99_s": pct(lat_f, 99), "p50_converged_s": pct(lat_f[conv], 50),
                 "p99_converged_s": pct(lat_f[conv], 99)},
        "served_over_fifo": (si_s / span_s) / (si_f / span_f),
        "same_final_hausdorff": bool(same_iters and h_diff == 0.0),
        "server": {k: info[k] for k in ("epochs", "rebuilds", "rebuilds_grow", "compactions", "extent_updates",
                                         "hausdorff_checks", "peak_bytes", "busy_seconds",
                                         "step_seconds", "check_seconds", "rebuild_seconds") if k in info},
    }


def bench_serve(args) -> int:
    """Serving form, printed as its own JSON line."""
    r = run_serve(args.arrivals, args.rate, args.ingest_threads, args.serve_log)
    out = {
        "metric": "served Lloyd site-iters/s on a Poisson request trace; p50/p99 time-to-convergence",
        "value": r["site_iters_per_s"], "unit": "site-iters/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded icospheres, tori, displaced spheres; Poisson arrivals)",
        "config": {"workload": "serve_trace", "requests": r["requests"], "rate_per_s": args.rate,
                   "large": r["large"], "K": 24, "freezing": True},
        "serve": r,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_check(args) -> int:
    """Latency of the convergence check (cvt_job_hausdorff_within) on a
    workload after --warmup iterations: the bounded verdict alone (a tolerance
    the job does not meet yet) and the verdict plus the exact distance (a
    tolerance it meets), each the median of --steps calls."""
    from gpu_cvt_amd import cvt, meshgen as mg

    V, F, S, K, _iters, _fz = mg.workload(args.workload)
    t0 = time.perf_counter()
    job = cvt.Job(V, F, S, cvt.default_params(K=K, freeze_mode=0))
    t_create = time.perf_counter() - t0
    job.step(args.warmup)
    _ok, h = job.hausdorff_within(1e30)  # builds the LBVH; the distance now
    out = {}
    for name, tol in [("beyond", 0.5 * h), ("within", 2.0 * h)]:
        ts = []
        for _ in range(max(1, args.steps)):
            t1 = time.perf_counter()
            ok, _h = job.hausdorff_within(tol)
            ts.append(time.perf_counter() - t1)
            assert ok == (name == "within")
        out[name] = 1e3 * sorted(ts)[len(ts) // 2]
        print(f"[bench] check {name} tolerance: median {out[name]:.2f} ms over {len(ts)} calls", file=sys.stderr,
              flush=True)
    job.close()
    print(json.dumps({"metric": "convergence check latency (p50 of cvt_job_hausdorff_within)",
                      "value": out["beyond"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": args.workload, "sites": len(S), "triangles": len(F), "K": K},
                      "within_with_distance_ms": out["within"], "job_create_ms": 1e3 * t_create,
                      "hausdorff": h}))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="dsphere_2m_1m")
    ap.add_argument("--freeze", action="store_true", help="freezing on (reported alongside)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--profile-iters", type=int, default=5)
    ap.add_argument("--mode", choices=["plain", "batch", "serve", "check"], default="plain",
                    help="plain: the headline 1M-site job; batch: 256 mixed jobs, batched vs one at a time; "
                         "serve: Poisson request trace through the serving loop vs FIFO plain path; "
                         "check: convergence-check latency on the workload")
    ap.add_argument("--arrivals", type=int, default=2000, help="serve mode: requests in the trace")
    ap.add_argument("--hausdorff-samples", type=int, default=0,
                    help="plain mode: after the timed steps, symmetric Hausdorff remesh<->input with this many "
                         "area-uniform samples per surface (0 = skip)")
    ap.add_argument("--rate", type=float, default=200.0, help="serve mode: arrivals per second")
    ap.add_argument("--ingest-threads", type=int, default=4, help="serve mode: server threads creating jobs")
    ap.add_argument("--serve-log", default="", help="serve mode: write per-request times (JSON) to this path")
    ap.add_argument("--jobs", type=int, default=256, help="batch mode: number of jobs")
    ap.add_argument("--site-cell", type=float, default=0.0, help="site grid cell scale (0 = library default)")
    ap.add_argument("--vert-cell", type=float, default=0.0, help="vertex grid cell scale (0 = library default)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (torch), to calibrate HBM counters")
    ap.add_argument("--gather-calib", type=int, default=0, metavar="MIB",
                    help="first run one gather of MIB MiB in 128-byte rows of random order (torch index_select; "
                         "every row read once), to calibrate HBM counters on a gather")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain mode: skip the freezing-on, K = 8, batched and served measurements")
    ap.add_argument("--extra-arrivals", type=int, default=400,
                    help="plain mode: requests of the served-form measurement folded into the line")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the remesher has no multi-GPU form "
              "(jobs would shard across GPUs with no communication between them).", file=sys.stderr)
        return 2

    from gpu_cvt_amd import cvt, meshgen as mg

    if args.copy_calib > 0:
        import torch

        n = args.copy_calib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {args.copy_calib} MiB read + {args.copy_calib} MiB written",
              file=sys.stderr, flush=True)
        del src, dst
    if args.gather_calib > 0:
        import torch

        rows = args.gather_calib * (1 << 20) // 128
        src = torch.ones((rows, 32), dtype=torch.float32, device="cuda")
        perm = torch.randperm(rows, device="cuda")
        torch.cuda.synchronize()
        dst = torch.index_select(src, 0, perm)
        torch.cuda.synchronize()
        print(f"[bench] gather calibration: {args.gather_calib} MiB read in {rows} random 128-byte rows "
              f"(+ {8 * rows} index bytes) + {args.gather_calib} MiB written", file=sys.stderr, flush=True)
        del src, dst, perm

    if args.mode == "batch":
        return bench_batch(args)
    if args.mode == "serve":
        return bench_serve(args)
    if args.mode == "check":
        return bench_check(args)

    t_gen = time.perf_counter()
    V, F, S, K, _iters, _fz = mg.workload(args.workload)
    t_gen = time.perf_counter() - t_gen
    params = cvt.default_params(K=K, freeze_mode=1 if args.freeze else 0, freeze_need=5, freeze_disp_rel=1e-5,
                                site_cell_scale=args.site_cell, vert_cell_scale=args.vert_cell)
    job = cvt.Job(V, F, S, params)
    info = job.info()
    print(f"[bench] workload={args.workload} nV={len(V)} nF={len(F)} nS={len(S)} K={K} gen={t_gen:.1f}s "
          f"site_grid={info['site_grid_dims']} cell={info['site_cell']:.4g} vert_grid={info['vert_grid_dims']} "
          f"dev_bytes={info['device_bytes'] / 2**20:.0f}MiB", file=sys.stderr, flush=True)

    job.step(args.warmup)
    stats = job.step(1, stats=True)[0]
    print(f"[bench] warm stats {stats}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms = job.step_timed(args.steps)
    wall = time.perf_counter() - t0
    ms_per_step = ms / args.steps
    value = len(S) * args.steps / (ms / 1e3)

    quality = None
    if args.hausdorff_samples > 0:
        # symmetric Hausdorff between the remeshed surface and the input mesh,
        # area-uniform samples on both, distances through device LBVHs
        T = job.remesh()
        Sf = job.sites()
        h = cvt.hausdorff(Sf, T, V, F, args.hausdorff_samples, seed=3, symmetric=True)
        h1 = cvt.hausdorff(Sf, T, V, F, args.hausdorff_samples, seed=3, symmetric=False)
        quality = {"iterations": args.warmup + 1 + args.steps, "remeshed_triangles": int(len(T)),
                   "hausdorff_symmetric": h, "hausdorff_remesh_to_input": h1,
                   "bbox_diag": info["bbox_diag"], "samples_per_surface": args.hausdorff_samples}
        print(f"[bench] quality {quality}", file=sys.stderr, flush=True)
    prof = job.profile(args.profile_iters)
    alg = algorithmic_bytes(len(S), len(V), len(F), K)
    achieved = alg / (prof["lloyd_ms"] * 1e-3) / 1e9
    traffic, traffic_src = traffic_from_profile(args.workload)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": "FP32 Lloyd pass: kb_cell<24> + kb_project"}
    # the roof that binds these kernels: their VALU issue floor (from the
    # committed SQ-counter profile) against the same live FP32-pass time
    valu_us = valu_issue_from_profile(args.workload)
    if valu_us is not None:
        roofline["valu_issue_us"] = valu_us
        roofline["valu_frac"] = round(valu_us / (prof["lloyd_ms"] * 1e3), 4)
    whole_iter_gbs = alg / (ms_per_step * 1e-3) / 1e9
    print(f"[bench] timed {args.steps} steps: {ms:.3f} ms device ({wall * 1e3:.1f} ms wall); "
          f"profile {prof}; kernel {achieved:.1f} GB/s; whole iteration {whole_iter_gbs:.1f} GB/s",
          file=sys.stderr, flush=True)
    job.close()

    extra = {}
    if not args.no_extra:
        # the same job with freezing on (the headline's second variant):
        # algorithmic bytes from each timed iteration's active and frozen sites
        fz = cvt.Job(V, F, S, cvt.default_params(K=K, freeze_mode=1, freeze_need=5, freeze_disp_rel=1e-5))
        fz.step(args.warmup)
        ms_f = fz.step_timed(args.steps)
        st = fz.last_stats(args.steps + 1)
        frozen_before = [s_["frozen_total"] for s_ in st[:-1]]
        fz.close()
        b_site = [(len(S) - f) * (29 + 8 * K) + f * (25 + 4 * K) for f in frozen_before]
        b_iter = float(np.mean(b_site)) + 28 * len(V) + 24 * len(F) + 4
        extra["freezing_on"] = {
            "site_iters_per_s": len(S) * args.steps / (ms_f / 1e3), "ms_per_step": ms_f / args.steps,
            "frozen_fraction_end": st[-1]["frozen_total"] / len(S), "algorithmic_bytes_per_iteration": b_iter,
            "whole_iteration_hbm_frac": round(b_iter / (ms_f / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
        print(f"[bench] freezing on: {extra['freezing_on']}", file=sys.stderr, flush=True)
        # K = 8 on the same mesh and sites: few neighbours leave many cells open
        # and pile sites onto the same mesh points, which the FP64 pass takes
        with cvt.Job(V, F, S, K=8) as j8:
            j8.step(args.warmup)
            ms_8 = j8.step_timed(args.steps)
            st8 = j8.last_stats(1)[0]
        extra["small_k"] = {"K": 8, "site_iters_per_s": len(S) * args.steps / (ms_8 / 1e3),
                            "ms_per_step": ms_8 / args.steps, "over_headline_ms": (ms_8 / args.steps) / ms_per_step,
                            "deferred_sites_last": st8["deferred_sites"]}
        print(f"[bench] K = 8: {extra['small_k']}", file=sys.stderr, flush=True)
        try:
            extra["batch"] = run_batch(256, args.steps, args.warmup)
        except Exception as e:
            print(f"[bench] batch measurement failed: {e!r}", file=sys.stderr)
        try:
            extra["serve"] = run_serve(args.extra_arrivals, 200.0)
        except Exception as e:
            print(f"[bench] serve measurement failed: {e!r}", file=sys.stderr)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(V, F, S, K, args.cpu_budget)
        except Exception as e:  # the CPU leg must not hide the GPU result
            print(f"[bench] cpu baseline failed: {e!r}", file=sys.stderr)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced sphere, area-uniform sites)",
        "config": {"workload": args.workload, "triangles": int(len(F)), "vertices": int(len(V)),
                   "sites": int(len(S)), "K": K, "freezing": bool(args.freeze), "kproj": 5},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "whole_iteration_hbm_frac": round(whole_iter_gbs / HBM_PEAK_GBS, 5),
        "algorithmic_bytes_per_launch": alg,
        "traffic_source": traffic_src,
        "stage_ms": prof,
        "quality": quality,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
