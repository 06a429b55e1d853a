"""Switches for the reference's variants, each against the FP64 oracle with
the same switch:

  * exclude_self_vertex: the command line's mode 0 drops mesh vertex i from
    site i's centroid -> mesh kproj-NN (knn_vertices_bruteforce_k<KPROJ, true>,
    main.cu:728, near_project.cuh:72-74); sites start at the mesh vertices;
  * kproj = 10: the Python implementation's projection list length
    (scripts/testfreeze.py:379), with frozen sites and with the FP64 pass;
  * remesh_R_mode 1: the remesh cells are clipped by the max bbox edge of the
    current sites, R2 (main.cu:899-909), not by the Lloyd R.

Lloyd steps are teacher-forced and checked stage by stage (lists bit-exact,
centroids and sites within 1e-5 of the bbox diagonal); the remesh triangle
sets must agree on all but co-circular cell vertices.
"""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg
from test_gpu_freeze import _force_mask, _teacher_forced_step

pytestmark = pytest.mark.gpu


def _thr(job):
    thr = 1e-5 * job.info()["bbox_diag"]
    return np.full(6, np.float32(thr * thr), dtype=np.float64), np.full(6, 5, np.int32)


@pytest.mark.parametrize("K", [24, 8])
def test_exclude_self_vertex(orc, K):
    V, F = mg.displaced_sphere(40, seed=51, crease_seed=52)
    mesh = orc.Mesh(V, F)
    S = mesh.V.copy()  # the command line starts the sites at the vertices
    print(f"[exclude_self] nV={len(V)} nS={len(S)} K={K}", flush=True)
    with cvt.Job(V, F, S, K=K, debug=1, exclude_self_vertex=1) as job, cvt.Job(V, F, S, K=K) as ref:
        thr2, need = _thr(job)
        job.step(1)  # the checked steps compare lists with a previous one
        for it in range(4):
            stats, ec, es, _ = _teacher_forced_step(orc, mesh, job, K, 0, thr2, need, exclude_self=True)
            print(f"  iter {it}: centroid err {ec:.2e}, site err {es:.2e} (x diag); stats {stats}", flush=True)
            kv = job.knnv()
            assert not (kv == np.arange(len(S))[:, None]).any()
        ref.step(5)
        # the switch matters: without it most lists hold the site's own vertex
        assert (ref.knnv() == np.arange(len(S))[:, None]).any(axis=1).mean() > 0.5


@pytest.mark.parametrize("K", [24, 8])
def test_kproj10(orc, K):
    V, F = mg.displaced_sphere(48, seed=53, crease_seed=54)
    mesh = orc.Mesh(V, F)
    S = mg.sample_surface(V, F, 8000, seed=55)
    print(f"[kproj10] nV={len(V)} nS={len(S)} K={K}", flush=True)
    with cvt.Job(V, F, S, K=K, kproj=10, debug=1, freeze_mode=1, freeze_need=5, freeze_disp_rel=1e-5) as job:
        thr2, need = _thr(job)
        job.step(3)
        _force_mask(job, np.random.default_rng(56), 0.1, 0.2, 5)
        for it in range(4):
            stats, ec, es, nfz = _teacher_forced_step(orc, mesh, job, K, 1, thr2, need, kproj=10)
            print(f"  iter {it}: frozen before {nfz}, centroid err {ec:.2e}, site err {es:.2e} (x diag); "
                  f"stats {stats}", flush=True)
            assert job.knnv().shape == (len(S), 10)


def test_remesh_R2_from_current_sites(orc):
    """Sites on a cap of the sphere (an open patch: its boundary cells are
    clipped by the square): with remesh_R_mode 1 the device matches the
    oracle's remesh at R2 = the sites' max bbox edge, whose triangle set
    differs from the one at the Lloyd R."""
    V, F = mg.icosphere(5)
    S = mg.sample_surface(V, F, 60000, seed=57)
    S = S[S[:, 2] > 0.95]
    K = 24
    print(f"[remesh R2] cap sites={len(S)}", flush=True)
    with cvt.Job(V, F, S, K=K, remesh_R_mode=1) as job:
        T = job.remesh()
        Sf = orc.f32r(job.sites())
    mesh = orc.Mesh(V, F)
    S32 = Sf.astype(np.float32)
    bb = S32.max(axis=0) - S32.min(axis=0)
    R2 = float(np.float32(2.0) * np.float32(0.5) * bb.max())
    R = float(np.float32(mesh.bbox_max_edge))
    knn = orc.knn(Sf, Sf, K, True)
    nn = orc.knn(mesh.V, Sf, 1, False)[:, 0]
    U, W = orc.frames(mesh.N[nn])
    ref2 = orc.remesh(Sf, U, W, knn, R2, mesh.N[nn])
    ref1 = orc.remesh(Sf, U, W, knn, R, mesh.N[nn])
    key = lambda T_: {tuple(t) for t in np.sort(T_, axis=1)}  # noqa: E731
    got, e2, e1 = key(T), key(ref2), key(ref1)
    j2 = len(got & e2) / max(len(got | e2), 1)
    j1 = len(e1 & e2) / max(len(e1 | e2), 1)
    print(f"[remesh R2] R2={R2:.4f} R={R:.4f}: device {len(got)} oracle(R2) {len(e2)} oracle(R) {len(e1)} "
          f"jaccard device/R2 {j2:.6f}, R/R2 {j1:.6f}", flush=True)
    assert R2 < 0.5 * R
    only_R = e1 - e2  # boundary triangles whose cell vertex lies between R2 and R
    assert only_R, "the clip size does not change this remesh: the test would not see the switch"
    assert not (only_R & got), "the device remesh kept triangles only the Lloyd R allows"
    assert j2 >= 0.999
