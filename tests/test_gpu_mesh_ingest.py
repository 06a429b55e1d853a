"""Mesh ingest on the device: compact_mesh (main.cu:240-268) as
cvt_mesh_compact, and the job set-up's face / site orderings (made on the
device) seen through the results they must not change."""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def np_compact(V, F):
    """Independent restatement: keep referenced vertices in order, renumber."""
    used = np.zeros(len(V), bool)
    used[F.reshape(-1)] = True
    remap = np.cumsum(used) - 1
    return V[used], remap[F]


def test_compact_drops_unreferenced_vertices():
    V, F = mg.icosphere(3)
    rng = np.random.default_rng(71)
    # interleave 500 unreferenced vertices and renumber the faces to match
    n = len(V) + 500
    pos = np.sort(rng.choice(n, len(V), replace=False))
    Vb = rng.normal(size=(n, 3))
    Vb[pos] = V
    Fb = pos[F].astype(np.int32)
    print(f"[compact] nV={n} nF={len(F)}", flush=True)
    Vc, Fc = cvt.mesh_compact(Vb, Fb)
    assert np.array_equal(Vc, V) and np.array_equal(Fc, F)       # closed form: the original mesh back
    Vr, Fr = np_compact(Vb, Fb)
    assert np.array_equal(Vc, Vr) and np.array_equal(Fc, Fr)


def test_compact_edge_cases():
    V = np.arange(15, dtype=np.float64).reshape(5, 3)
    F = np.array([[4, 4, 2]], np.int32)                         # a repeated corner; 0, 1, 3 unused
    Vc, Fc = cvt.mesh_compact(V, F)
    assert np.array_equal(Vc, V[[2, 4]]) and np.array_equal(Fc, [[1, 1, 0]])
    Vc, Fc = cvt.mesh_compact(V, np.zeros((0, 3), np.int32))    # no faces: no vertex is kept
    assert len(Vc) == 0
    with pytest.raises(cvt.CVTError):
        cvt.mesh_compact(V, np.array([[0, 1, 5]], np.int32))


def test_site_order_only_renumbers(orc):
    """The set-up's site ordering (by site-grid cell, caller's index within a
    cell) only renumbers: a job made from shuffled sites gives the same sites
    and lists, in the caller's ids, as the unshuffled one after several
    iterations; and a step of it matches the oracle stage by stage."""
    from test_gpu_freeze import _teacher_forced_step
    V, F = mg.displaced_sphere(48, seed=72, crease_seed=73)
    S = mg.sample_surface(V, F, 12000, seed=74)
    ps = np.random.default_rng(75).permutation(len(S))
    inv_s = np.argsort(ps)
    print(f"[site order] nV={len(V)} nS={len(S)}", flush=True)
    with cvt.Job(V, F, S, K=24) as a, cvt.Job(V, F, S[ps], K=24, debug=1) as b:
        a.step(6)
        b.step(6)
        assert np.array_equal(a.sites(), b.sites()[inv_s])
        assert np.array_equal(a.knn(), ps[b.knn()[inv_s]])
        assert np.array_equal(a.knnv(), b.knnv()[inv_s])
        thr = 1e-5 * b.info()["bbox_diag"]
        thr2 = np.full(6, np.float32(thr * thr), dtype=np.float64)
        _teacher_forced_step(orc, orc.Mesh(V, F), b, 24, 0, thr2, np.full(6, 5, np.int32))
