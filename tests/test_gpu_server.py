"""Serving form: requests of different sizes and meshes, admitted under an HBM
budget and advanced together in epochs, retire at convergence or max_iters.
Each request's final sites must be bit-identical to the same job run alone for
the same number of iterations, whatever else shared its epochs."""
import time

import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def _requests():
    out = []
    V, F = mg.icosphere(4)
    out.append((V, F, mg.sample_surface(V, F, 800, seed=31), dict(freeze_mode=1, freeze_disp_rel=1e-4), 2e-3, 40))
    V, F = mg.torus(1.0, 0.3, 100, 40, jitter=1e-3, seed=32)
    out.append((V, F, mg.sample_surface(V, F, 3000, seed=33), dict(), 1e-9, 30))   # never converges: max_iters
    V, F = mg.displaced_sphere(25, seed=34, crease_seed=35)
    out.append((V, F, mg.sample_surface(V, F, 2500, seed=36), dict(freeze_mode=1, freeze_disp_rel=1e-3), 5e-3, 60))
    V, F = mg.icosphere(3)
    out.append((V, F, mg.sample_surface(V, F, 300, seed=37), dict(), 1e-2, 20))
    return out


def _solo(V, F, S, kw, iters):
    with cvt.Job(V, F, S, K=24, **kw) as j:
        j.step(iters)
        return j.sites()


@pytest.mark.parametrize("budget", [None, 1], ids=["budget_200g", "one_at_a_time"])
def test_server_results_match_solo(budget):
    reqs = _requests()
    print(f"[server] budget={budget} requests={[len(r[2]) for r in reqs]}", flush=True)
    with cvt.Server(hbm_budget_bytes=budget) as srv:
        tickets = []
        for k, (V, F, S, kw, tol, mx) in enumerate(reqs):
            tickets.append(srv.submit(V, F, S, tol_rel=tol, max_iters=mx, K=24, **kw))
            if k == 1:
                time.sleep(0.05)   # later arrivals join a running batch
        srv.wait_idle(300)
        info = srv.info()
        print(f"[server] info {info}", flush=True)
        assert info["failures"] == 0 and info["active"] == 0 and info["pending"] == 0
        # one slot batch for the (K, kproj) group, captured once: jobs join and
        # retire by editing its slot table in place
        assert info["rebuilds"] == 1
        for k, t in enumerate(tickets):
            r = srv.poll(t)
            V, F, S, kw, tol, mx = reqs[k]
            print(f"[server] request {k}: {r}", flush=True)
            assert r["state"] in ("converged", "max_iters")
            assert r["iterations"] % srv.config.epoch_iters == 0 and r["iterations"] <= mx
            assert r["t_arrive"] <= r["t_admit"] <= r["t_done"]
            if r["state"] == "converged":
                assert 0 <= r["hausdorff"] <= tol * r["bbox_diag"]
            else:
                assert r["iterations"] == mx
            got = srv.sites(t)
            ref = _solo(V, F, S, kw, r["iterations"])
            assert np.array_equal(got, ref), f"request {k}: served result differs from the solo run"
        assert srv.poll(tickets[1])["state"] == "max_iters"


def test_server_slot_table_grows_and_reuses_slots():
    """Twenty requests in three waves: jobs join a running batch and retired
    jobs' slots and block ranges are reused in place, without capturing the
    batch graph again; the launch extents follow the tables in use."""
    V, F = mg.icosphere(3)
    reqs = []
    for k in range(20):
        n = 200 + 150 * (k % 5)
        kw = dict(freeze_mode=1, freeze_disp_rel=1e-4) if k % 3 == 0 else dict()
        reqs.append((V, F, mg.sample_surface(V, F, n, seed=300 + k), kw, 10 * (1 + k % 4)))
    _serve_and_compare(reqs, {5: 0.2, 13: 0.2}, dict(), lambda info: info["rebuilds"] == 1)


def test_server_batch_tables_grow_and_compact():
    """Small first capacities (4 jobs, 512 site blocks).  A 60k-site job runs
    alone, then five small jobs join after it (the slot table outgrows 4 jobs:
    the tables grow and the batch graph is captured again); the large job
    retires first, leaving a hole below the small jobs' block ranges (the
    tables are compacted in place); ten more small jobs join.  Every result is
    bit-identical to its solo run."""
    Vb, Fb = mg.displaced_sphere(128, seed=41, crease_seed=42)
    V, F = mg.icosphere(3)
    # iteration counts with margin on both sides of the race: the large job
    # must still run when the five small jobs join (it has done >= 10 of its
    # 200 when they are submitted) and retire before they do (400)
    reqs = [(Vb, Fb, mg.sample_surface(Vb, Fb, 60_000, seed=43), dict(), 200)]
    for k in range(15):
        reqs.append((V, F, mg.sample_surface(V, F, 300 + 40 * k, seed=500 + k), dict(), 400 if k < 5 else 30))

    def big_running(srv, tickets):
        t_end = time.time() + 120
        while srv.poll(tickets[0])["iterations"] < 10:
            assert time.time() < t_end, "the first request never ran"
            time.sleep(0.005)

    def check(info):
        return info["rebuilds_grow"] >= 1 and info["compactions"] >= 1 and info["extent_updates"] >= 2

    _serve_and_compare(reqs, {0: big_running, 5: 0.3}, dict(batch_jobs0=4, batch_site_blocks0=512), check)


def _serve_and_compare(reqs, pause_after, cfg, info_ok):
    print(f"[server-slots] cfg={cfg} sizes={[len(r[2]) for r in reqs]}", flush=True)
    with cvt.Server(**cfg) as srv:
        tickets = []
        for k, (V, F, S, kw, mx) in enumerate(reqs):
            tickets.append(srv.submit(V, F, S, tol_rel=1e-9, max_iters=mx, K=24, **kw))
            pause = pause_after.get(k)
            if callable(pause):
                pause(srv, tickets)
            elif pause:
                time.sleep(pause)
        srv.wait_idle(300)
        info = srv.info()
        print(f"[server-slots] info {info}", flush=True)
        assert info["failures"] == 0 and info["active"] == 0
        assert info_ok(info), info
        for k, t in enumerate(tickets):
            r = srv.poll(t)
            V, F, S, kw, mx = reqs[k]
            assert r["state"] == "max_iters" and r["iterations"] == mx, (k, r)
            assert np.array_equal(srv.sites(t), _solo(V, F, S, kw, mx)), f"request {k} differs from its solo run"
