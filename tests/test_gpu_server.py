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
        if budget == 1:
            assert info["rebuilds"] >= len(reqs)
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
