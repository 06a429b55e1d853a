"""Freezing against the FP64 oracle at the second configuration's full size
(jittered torus, 200,000 triangles, 100,000 sites, K = 24), and one
teacher-forced iteration of the 1M-site headline job.

Every iteration is teacher-forced: the GPU's state before the step (FP32
sites, previous and pinned lists, frozen mask, streaks, tiers) is handed to
the oracle's full Lloyd step (orc.lloyd_step, main.cu:669-881 restated), and
every stage of every site -- frozen ones included -- is compared:
  * site KNN lists (frozen rows: the pinned lists), bit-exact;
  * the frame vertex of every site (frozen: the pinned vertex list's first
    entry, main.cu:676-724), bit-exact;
  * centroids and projected sites, within 1e-5 of the bbox diagonal;
  * centroid -> mesh vertex lists, bit-exact, from the GPU's FP32 centroids
    (a list depends on which FP32 rounding of the centroid it is taken at);
  * streaks, frozen flags and the same/low/both/newly-frozen counters, exact,
    from the GPU's FP32 sites before and after the step
    (freezetest.cuh:155-212,270-327,451-466).
"""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu

TOL_REL_DIAG = 1e-5


def _teacher_forced_step(orc, mesh, job, K, freeze_mode, thr2, need, tier=None, kproj=5, exclude_self=False):
    """One GPU step from the job's current state, checked stage by stage
    against orc.lloyd_step from the same state.  Returns the GPU's stats and
    the largest centroid / site errors (as fractions of the diagonal)."""
    diag = mesh.bbox_diag
    R = np.float32(mesh.bbox_max_edge)
    n = job.nS
    S32 = orc.f32r(job.sites())
    knn_prev = job.knn()
    knnv_prev = job.knnv()
    fz0, st0 = job.freeze_state()
    st = orc.LloydState(S32, K, kproj=kproj)
    st.knn[:] = knn_prev
    st.knnV[:] = knnv_prev
    st.frozen[:] = fz0
    st.streak[:] = st0
    st.has_prev = 1
    if tier is not None:
        st.tier[:] = tier
    stats = job.step(1, stats=True)[0]
    out = orc.lloyd_step(mesh, st, K=K, kproj=kproj, R=R, freeze_mode=freeze_mode, thresh2_tier=thr2,
                         streak_tier=need, exclude_self=exclude_self)
    knn_g = job.knn()
    knnv_g = job.knnv()
    dbg = job.debug()
    S_new = job.sites()
    fz1, st1 = job.freeze_state()
    frozen = fz0 == 1
    # 1. site KNN (frozen rows keep their pinned lists)
    bad = np.any(knn_g != out["knn"], axis=1)
    assert not bad.any(), f"site KNN differs in {bad.sum()} rows ({np.sum(bad & frozen)} frozen)"
    assert np.array_equal(knn_g[frozen], knn_prev[frozen])
    # 2. frame vertex (frozen sites: the pinned vertex list's first entry)
    assert np.array_equal(dbg["nn_site"], out["nn_site"]), \
        f"frame vertex differs at {np.sum(dbg['nn_site'] != out['nn_site'])} sites"
    assert np.array_equal(dbg["nn_site"][frozen], knnv_prev[frozen, 0])
    # 3. centroids
    ec = np.linalg.norm(dbg["cent"] - out["cent"], axis=1)
    assert ec.max() <= TOL_REL_DIAG * diag, f"centroid error {ec.max() / diag:.2e} x diag (frozen max " \
                                            f"{ec[frozen].max() / diag if frozen.any() else 0:.2e})"
    # 4. vertex lists: active sites from the GPU's FP32 centroids, frozen sites pinned
    cent32 = orc.f32r(dbg["cent"])
    act = ~frozen & np.all(np.isfinite(cent32), axis=1)
    knnv_o = knnv_prev.copy()
    knnv_o[~frozen] = -1
    if exclude_self:  # query i skips vertex i: search every site in order
        knnv_o[act] = orc.knn(mesh.V, np.where(np.isfinite(cent32), cent32, 0.0), kproj, exclude_self=True)[act]
    else:
        knnv_o[act] = orc.knn(mesh.V, cent32[act], kproj, exclude_self=False)
    assert np.array_equal(knnv_g, knnv_o), f"vertex lists differ in {np.sum(np.any(knnv_g != knnv_o, 1))} rows"
    assert np.array_equal(out["knnV"][frozen], knnv_prev[frozen])
    # 5. projected sites: against the full oracle step, and against the
    #    projection of the GPU's own centroid over the same lists
    es = np.linalg.norm(S_new - out["Snew"], axis=1)
    assert es.max() <= TOL_REL_DIAG * diag, f"site error {es.max() / diag:.2e} x diag"
    proj = orc.project(cent32, knnv_o, mesh.V, mesh.F, mesh.vf)
    ok = np.all(np.isfinite(proj), axis=1)
    ep = np.linalg.norm(S_new[ok] - proj[ok], axis=1)
    assert ep.max() <= TOL_REL_DIAG * diag
    # 6. freeze state machine on the GPU's FP32 sites before and after
    exp_st, exp_fz, counts = orc.freeze(S32, orc.f32r(S_new), knn_g, knn_prev, tier, thr2, need, 1,
                                        int(freeze_mode != 0), st0, fz0)
    assert np.array_equal(st1, exp_st), f"streak differs at {np.sum(st1 != exp_st)} sites"
    assert np.array_equal(fz1, exp_fz), f"frozen flag differs at {np.sum(fz1 != exp_fz)} sites"
    assert (stats["count_same"], stats["count_low_disp"], stats["count_both"], stats["count_newly_frozen"]) == \
        tuple(int(c) for c in counts)
    assert stats["frozen_total"] == int(fz1.sum())
    return stats, ec.max() / diag, es.max() / diag, int(frozen.sum())


def _torus_job(mode, **kw):
    V, F, S, K, _iters, _fz = mg.workload("torus_200k_100k")
    return V, F, S, K, cvt.Job(V, F, S, K=K, debug=1, freeze_mode=mode, **kw)


def _force_mask(job, rng, frac_frozen, frac_primed, need_of):
    """Freeze a random fraction of the sites at their current lists and prime
    another fraction one step short of freezing (so the iteration freezes the
    ones whose lists and displacements pass)."""
    n = job.nS
    fz, st = job.freeze_state()
    pick = rng.random(n)
    fz = np.where(pick < frac_frozen, 1, fz).astype(np.uint8)
    primed = (pick >= frac_frozen) & (pick < frac_frozen + frac_primed)
    st = np.where(primed, np.maximum(need_of - 2, 0), st).astype(np.uint8)
    job.set_state(frozen=fz, streak=st)
    return fz


def test_freeze_mode1_torus_200k_100k(orc):
    """freeze_mode 1 (single-tier streak, freezetest.cuh:155-212) with the
    configuration's threshold of 1e-5 of the bbox diagonal."""
    V, F, S, K, job = _torus_job(1, freeze_need=5, freeze_disp_rel=1e-5)
    mesh = orc.Mesh(V, F)
    thr = 1e-5 * job.info()["bbox_diag"]
    thr2 = np.full(6, np.float32(thr * thr), dtype=np.float64)
    need = np.full(6, 5, np.int32)
    rng = np.random.default_rng(21)
    print(f"[freeze1] nV={len(V)} nF={len(F)} nS={len(S)} K={K}", flush=True)
    with job:
        job.step(8)
        _force_mask(job, rng, 0.12, 0.2, 5)
        for it in range(6):
            stats, ec, es, nfz = _teacher_forced_step(orc, mesh, job, K, 1, thr2, need)
            print(f"  iter {it}: frozen before {nfz}, centroid err {ec:.2e}, site err {es:.2e} (x diag); "
                  f"stats {stats}", flush=True)
            assert nfz >= 0.1 * len(S)


@pytest.mark.parametrize("shape", ["torus_200k_100k", "cube_6k"])
def test_freeze_mode2_device_tiers(orc, shape):
    """freeze_mode 2 with per-site tiers computed on the device (scheme 2,
    main.cu:41-149) and per-tier streaks and thresholds (freezetest.cuh:270-327),
    larger thresholds on flatter tiers so that sites freeze during the run.
    The smooth torus of the second configuration is all in the flattest tiers;
    the cube's edges and corners reach the higher ones."""
    if shape == "cube_6k":
        V, F = mg.cube(24)
        S = mg.sample_surface(V, F, 6000, seed=6)
        K = 24
    else:
        V, F, S, K, _iters, _fz = mg.workload(shape)
    mesh = orc.Mesh(V, F)
    base = 2e-4 * mesh.bbox_diag
    thr_lin = base * np.array([4.0, 2.0, 1.0, 0.5, 0.25, 0.25])
    thr2 = (thr_lin ** 2).astype(np.float32).astype(np.float64)
    need = np.array([3, 4, 5, 6, 7, 4], np.int32)
    job = cvt.Job(V, F, S, K=K, debug=1, freeze_mode=2, tier_streak=need.tolist(), tier_disp_abs2=thr2.tolist())
    rng = np.random.default_rng(22)
    print(f"[freeze2] {shape}: nV={len(V)} nF={len(F)} nS={len(S)} K={K}", flush=True)
    with job:
        job.step(8)
        tier, _nv, _L = job.compute_tiers(2)
        hist = np.bincount(tier, minlength=6)
        print(f"[freeze2] tiers {hist.tolist()}", flush=True)
        if shape == "cube_6k":
            assert np.count_nonzero(hist) >= 3
        _force_mask(job, rng, 0.12, 0.2, need[tier])
        newly = 0
        for it in range(6):
            stats, ec, es, nfz = _teacher_forced_step(orc, mesh, job, K, 2, thr2, need, tier=tier)
            newly += stats["count_newly_frozen"]
            print(f"  iter {it}: frozen before {nfz}, centroid err {ec:.2e}, site err {es:.2e} (x diag); "
                  f"stats {stats}", flush=True)
            assert nfz >= 0.1 * len(S)
        assert newly > 0, "no site froze during the checked iterations"


def test_headline_job_iteration_20(orc):
    """The 1M-site headline job (displaced sphere, 1,997,568 triangles,
    K = 24, freezing off): 20 free-running iterations, then one teacher-forced
    iteration checked stage by stage against the oracle."""
    V, F, S, K, _iters, _fz = mg.workload("dsphere_2m_1m")
    mesh = orc.Mesh(V, F)
    print(f"[headline] nV={len(V)} nF={len(F)} nS={len(S)} K={K}", flush=True)
    need = np.full(6, 5, np.int32)
    with cvt.Job(V, F, S, K=K, debug=1) as job:
        # freezing off still counts low displacements against the default threshold (1e-5 x diag)
        thr = 1e-5 * job.info()["bbox_diag"]
        thr2 = np.full(6, np.float32(thr * thr), dtype=np.float64)
        job.step(20)
        stats, ec, es, _ = _teacher_forced_step(orc, mesh, job, K, 0, thr2, need)
        print(f"  iteration 21: centroid err {ec:.2e}, site err {es:.2e} (x diag); stats {stats}", flush=True)
