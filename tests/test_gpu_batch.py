"""Batched form: every job's result in a batch is bit-identical to running it
alone (sites, KNN lists, freeze state), including when batch steps and solo
steps are interleaved."""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def _inputs():
    out = []
    V, F = mg.icosphere(3)
    out.append((V, F, mg.sample_surface(V, F, 300, seed=1), dict()))
    V, F = mg.icosphere(5)
    out.append((V, F, mg.sample_surface(V, F, 2000, seed=2), dict(freeze_mode=1, freeze_disp_rel=1e-3)))
    V, F = mg.torus(1.0, 0.3, 120, 50, jitter=1e-3, seed=3)
    out.append((V, F, mg.sample_surface(V, F, 6000, seed=4), dict()))
    V, F = mg.displaced_sphere(30, seed=5, crease_seed=6)
    out.append((V, F, mg.sample_surface(V, F, 4000, seed=7), dict(freeze_mode=1, freeze_disp_rel=1e-3)))
    V, F = mg.icosphere(2)
    out.append((V, F, mg.sample_surface(V, F, 40, seed=8), dict()))
    return out


def _state(job):
    fz, st = job.freeze_state()
    return job.sites(), job.knn(), fz, st


@pytest.mark.parametrize("schedule", ["batch4", "mixed"])
def test_batch_matches_solo(schedule):
    ins = _inputs()
    print(f"[batch] schedule={schedule} jobs={len(ins)} sites={[len(x[2]) for x in ins]}", flush=True)
    solo = [cvt.Job(V, F, S, K=24, **kw) for V, F, S, kw in ins]
    for j in solo:
        j.step(4)
    ref = [_state(j) for j in solo]
    for j in solo:
        j.close()
    jobs = [cvt.Job(V, F, S, K=24, **kw) for V, F, S, kw in ins]
    with cvt.Batch(jobs) as b:
        if schedule == "batch4":
            b.step(4)
        else:
            b.step(1)
            jobs[1].step(1)          # one job advances alone in between
            jobs[3].step(1)
            b.step(1)                # jobs 1 and 3 are now one iteration ahead ...
            for k in (0, 2, 4):
                jobs[k].step(1)      # ... the others catch up alone
            for k in (0, 2, 4):
                jobs[k].step(1)
            for k in (1, 3):
                jobs[k].step(1)
    for k, j in enumerate(jobs):
        got = _state(j)
        for a, r, name in zip(got, ref[k], ("sites", "knn", "frozen", "streak")):
            assert np.array_equal(a, r), f"job {k}: {name} differs between batch and solo"
        j.close()


def test_batch_rejects_mixed_k():
    V, F = mg.icosphere(3)
    a = cvt.Job(V, F, mg.sample_surface(V, F, 200, 1), K=24)
    b = cvt.Job(V, F, mg.sample_surface(V, F, 200, 2), K=12)
    with pytest.raises(cvt.CVTError):
        cvt.Batch([a, b])
    a.close()
    b.close()


def test_batch_of_32_workload_jobs_matches_solo():
    """32 jobs of the batched-form distribution (icospheres, tori, displaced
    spheres; 5k-400k triangles, 1k-100k sites; K = 24, freezing on): the batch
    gives every job the state it reaches alone."""
    work = mg.batch_workload(32, seed=7)
    kw = dict(K=24, freeze_mode=1, freeze_need=5, freeze_disp_rel=1e-5)
    print(f"[batch32] sites {sum(len(w[2]) for w in work)}", flush=True)
    ref = []
    for V, F, S, _tag in work:
        with cvt.Job(V, F, S, **kw) as j:
            j.step(12)
            ref.append(_state(j))
    jobs = [cvt.Job(V, F, S, **kw) for V, F, S, _tag in work]
    try:
        with cvt.Batch(jobs) as b:
            b.step(12)
        for k, (j, r) in enumerate(zip(jobs, ref)):
            got = _state(j)
            for a, e in zip(got, r):
                assert np.array_equal(a, e), f"job {k} ({work[k][3]}) differs from its solo run"
    finally:
        for j in jobs:
            j.close()


def test_fp64_pass_runs_are_deterministic():
    """Jobs whose FP64 pass is busy (K = 8: open cells, far centroids, piles of
    coincident sites; it runs beside the projection kernel): two solo runs and
    a batched run reach bit-identical states."""
    V, F = mg.icosphere(4)
    ins = [(V, F, mg.sample_surface(V, F, 1500, seed=28)), (V, F, mg.sample_surface(V, F, 40000, seed=29))]
    V6, F6 = mg.icosphere(6)
    ins.append((V6, F6, mg.sample_surface(V6, F6, 40960, seed=30)))

    def run():
        out = []
        for V_, F_, S_ in ins:
            with cvt.Job(V_, F_, S_, K=8, freeze_mode=1, freeze_disp_rel=1e-4) as j:
                st = j.step(15, stats=True)
                out.append((_state(j), sum(s["deferred_sites"] for s in st)))
        return out

    a, b = run(), run()
    print(f"[determinism] deferred sites over 15 iterations: {[x[1] for x in a]}", flush=True)
    assert any(x[1] > 0 for x in a)
    for k, (x, y) in enumerate(zip(a, b)):
        for p, q in zip(x[0], y[0]):
            assert np.array_equal(p, q), f"job {k}: two solo runs differ"
    jobs = [cvt.Job(V_, F_, S_, K=8, freeze_mode=1, freeze_disp_rel=1e-4) for V_, F_, S_ in ins]
    try:
        with cvt.Batch(jobs) as bt:
            bt.step(15)
        for k, (j, x) in enumerate(zip(jobs, a)):
            for p, q in zip(_state(j), x[0]):
                assert np.array_equal(p, q), f"job {k}: batched run differs from solo"
    finally:
        for j in jobs:
            j.close()
