"""Remesh extraction and the device Hausdorff check against the oracle.

The device labels cell vertices from FP32 dual hulls; the oracle clips the
same cells in FP64 (Sutherland-Hodgman, 2dClip.cuh:32-66, labels as in
buildmesh.cuh:170-241).  Only cell vertices where four sites are nearly
co-circular can be labelled differently, so the triangle sets must agree on
all but a tiny fraction.  The Hausdorff value (remesh -> input, over face
centroids and edge midpoints, evaluate.cuh:252-283) must agree within 0.1%.
"""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def _cases():
    V, F = mg.icosphere(5)
    yield "ico5_3k_k24", V, F, mg.sample_surface(V, F, 3000, seed=21), 24
    V, F = mg.torus(1.0, 0.35, 160, 60, jitter=1e-3, seed=22)
    yield "torus_8k_k20", V, F, mg.sample_surface(V, F, 8000, seed=23), 20
    V, F = mg.displaced_sphere(40, seed=24, crease_seed=25)
    yield "dsphere_6k_k32", V, F, mg.sample_surface(V, F, 6000, seed=26), 32


def _oracle_remesh(orc, job, V, F, S, K):
    mesh = orc.Mesh(V, F)
    Sf = orc.f32r(S)
    knn = orc.knn(Sf, Sf, K, True)
    nn = orc.knn(mesh.V, Sf, 1, False)[:, 0]
    U, W = orc.frames(mesh.N[nn])
    R = float(np.float32(mesh.bbox_max_edge))
    return orc.remesh(Sf, U, W, knn, R, mesh.N[nn]), mesh


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_remesh_and_hausdorff_match_oracle(orc, case):
    name, V, F, S0, K = case
    print(f"[remesh] {name} sites={len(S0)} K={K}", flush=True)
    job = cvt.Job(V, F, S0, K=K)
    job.step(5)
    S = job.sites()
    T = job.remesh()
    ref, mesh = _oracle_remesh(orc, job, V, F, S, K)
    got = {tuple(t) for t in np.sort(T, axis=1)}
    exp = {tuple(t) for t in np.sort(ref, axis=1)}
    inter = len(got & exp)
    jacc = inter / max(len(got | exp), 1)
    print(f"[remesh] {name}: device {len(got)} oracle {len(exp)} common {inter} jaccard {jacc:.6f}", flush=True)
    assert jacc >= 0.999
    # orientation agrees on the common triangles (same cyclic order)
    cyc = lambda t: min((t[0], t[1], t[2]), (t[1], t[2], t[0]), (t[2], t[0], t[1])) # noqa: E731
    og = {tuple(sorted(t)): cyc(tuple(t)) for t in T}
    oe = {tuple(sorted(t)): cyc(tuple(t)) for t in ref}
    common = [k for k in og if k in oe]
    same = sum(og[k] == oe[k] for k in common)
    assert same >= 0.999 * len(common)

    h_dev, nf = job.hausdorff()
    assert nf == len(T)
    P = S[T]                                                     # (m, 3, 3)
    probes = np.concatenate([P.mean(axis=1), 0.5 * (P[:, 0] + P[:, 1]), 0.5 * (P[:, 1] + P[:, 2]),
                             0.5 * (P[:, 2] + P[:, 0])])
    h_ref = float(np.sqrt(orc.point_mesh_d2(probes, mesh.V, mesh.F).max()))
    print(f"[remesh] {name}: hausdorff device {h_dev:.6e} oracle {h_ref:.6e}", flush=True)
    assert abs(h_dev - h_ref) <= 1e-3 * h_ref + 1e-7 * mesh.bbox_diag
    job.close()


def test_hausdorff_within_agrees_with_exact():
    """The bounded convergence test says "within" exactly when the exact
    one-sided distance is <= tol, and then reports that same distance."""
    V, F = mg.displaced_sphere(30, seed=24, crease_seed=25)
    with cvt.Job(V, F, mg.sample_surface(V, F, 5000, seed=27), K=24) as job:
        job.step(6)
        h, _ = job.hausdorff()
        for tol in (0.5 * h, np.nextafter(h, 0), h, 1.5 * h):
            w, hw = job.hausdorff_within(tol)
            print(f"[within] h={h:.9e} tol={tol:.9e} within={w} h_within={hw:.9e}", flush=True)
            assert w == (h <= tol)
            assert (hw == h) if w else (hw == -1.0)
