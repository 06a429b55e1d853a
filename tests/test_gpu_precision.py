"""Precision form at full size: displaced sphere (1,002,252 triangles, 8
spherical-harmonic bands l <= 16, 12 crease clamps), 500,000 sites, K = 32.

FP32 bulk with compensated area/moment sums; FP64 only for ill-conditioned
frames and near-degenerate predicates.  Each iteration is teacher-forced: the
FP64 oracle recomputes KNN, frames, cells and projection from the GPU's own
FP32 sites and centroids, and the errors are printed per iteration.  Bars:
KNN lists exact; centroid and projection errors <= 1e-5 of the bbox
diagonal; final CVT energy and one-sided Hausdorff distance within 0.1% of the
oracle's.  Freezing is off here so that every site runs every stage.
"""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def test_precision_form_dsphere_1m_500k(orc):
    V, F, S, K, _iters, _fz = mg.workload("dsphere_1m_500k")
    print(f"[precision] nV={len(V)} nF={len(F)} nS={len(S)} K={K}", flush=True)
    mesh = orc.Mesh(V, F)
    diag = mesh.bbox_diag
    R = np.float32(mesh.bbox_max_edge)
    rows = []
    with cvt.Job(V, F, S, K=K, debug=1) as job:
        for it in range(5):
            S32 = orc.f32r(job.sites())
            st = job.step(1, stats=True)[0]
            knn_g = job.knn()
            dbg = job.debug()
            knn_o = orc.knn(S32, S32, K, exclude_self=True)
            bad = int(np.sum(np.any(knn_g != knn_o, axis=1)))
            nn_o = orc.knn(mesh.V, S32, 1, exclude_self=False)[:, 0]
            U, W = orc.frames(mesh.N[nn_o])
            cent_o, area_o, mom_o, _ = orc.cells(S32, U, W, knn_o, R)
            ec = np.linalg.norm(dbg["cent"] - cent_o, axis=1)
            cent32 = orc.f32r(dbg["cent"])
            knnv_o = orc.knn(mesh.V, cent32, 5, exclude_self=False)
            proj_o = orc.project(cent32, knnv_o, mesh.V, mesh.F, mesh.vf)
            ep = np.linalg.norm(job.sites() - proj_o, axis=1)
            rows.append((it, bad, ec.max() / diag, ec.mean() / diag, ep.max() / diag, st["deferred_sites"]))
            print(f"[precision] iter {it}: knn rows differing {bad}; centroid err max {ec.max() / diag:.2e} "
                  f"mean {ec.mean() / diag:.2e} (x diag); projection err max {ep.max() / diag:.2e}; "
                  f"deferred {st['deferred_sites']}", flush=True)
            assert bad == 0
            assert ec.max() <= 1e-5 * diag and ep.max() <= 1e-5 * diag
        # final energy: GPU eval pass vs the oracle's cells at the same FP32 sites
        S32 = orc.f32r(job.sites())
        E_g = job.energy()
        knn_o = orc.knn(S32, S32, K, exclude_self=True)
        nn_o = orc.knn(mesh.V, S32, 1, exclude_self=False)[:, 0]
        U, W = orc.frames(mesh.N[nn_o])
        _c, _a, mom_o, _ = orc.cells(S32, U, W, knn_o, R)
        E_o = float(np.sum(mom_o))
        # final one-sided Hausdorff: the GPU's remeshed surface, its probes, distances in FP64
        h_g, nf = job.hausdorff()
        T = job.remesh()
        P = S32[T]
        probes = np.concatenate([P.mean(axis=1), 0.5 * (P[:, 0] + P[:, 1]), 0.5 * (P[:, 1] + P[:, 2]),
                                 0.5 * (P[:, 2] + P[:, 0])])
        h_o = float(np.sqrt(orc.point_mesh_d2(probes, mesh.V, mesh.F).max()))
    e_rel = abs(E_g - E_o) / E_o
    h_rel = abs(h_g - h_o) / h_o
    print(f"[precision] final energy GPU {E_g:.9e} oracle {E_o:.9e} rel err {e_rel:.2e}; "
          f"Hausdorff GPU {h_g:.6e} oracle {h_o:.6e} rel err {h_rel:.2e} ({nf} remeshed triangles)", flush=True)
    assert e_rel <= 1e-3
    assert h_rel <= 1e-3
