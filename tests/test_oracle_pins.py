"""Pins for the FP64 oracle (CPU only).

The reference ships no tests, meshes or golden vectors, so each oracle
function is pinned by checks that a plausible mistake would fail: closed-form
answers, brute force written independently (numpy), invariants and symmetry.

  function (oracle/cvt_oracle.cpp)   pinned by
  orc_knn_brute / orc_knn_grid       numpy lexsort brute force (independent),
                                     lattice exact ties, grid == brute
  orc_frames                         closed forms at +z, +x; orthonormality;
                                     switch exactly at |n_z| = 0.9
  orc_cells / orc_cell_poly          square lattice (a*a, a^4/6), hexagonal
                                     lattice ((sqrt3/2)a^2, 5 sqrt3/8 s^4),
                                     single neighbour rectangle, empty list,
                                     icosahedron fixed point
  orc_closest_point_tri              independent min over plane/edges/vertices
  orc_project                        sphere: radial projection bound
  orc_vertex_normals                 icosphere radial bound, flat plane exact
  orc_freeze                         hand-worked streak scenarios, 255
                                     saturation, <= boundary, absorbing frozen
  orc_point_mesh_d2                  brute == BVH; translated plane == t^2
  orc_quality                        closed forms: equilateral (Q 1, 60 deg),
                                     right isosceles (Q sqrt3 (sqrt2 - 1), 45),
                                     obtuse isosceles (25/25/130 shares),
                                     collinear (Q 0, 0/0/180); means of a
                                     mixed set; rotation and scale invariance
  orc_lloyd_step                     energy decreases; fixed point; frozen
                                     lists pinned; exclude_self lists == numpy
                                     brute force with vertex i removed; kproj
                                     10 lists == numpy brute force and the
                                     projection == the global nearest point
                                     (icosphere, independent brute force)
"""
import numpy as np
import pytest

from gpu_cvt_amd import meshgen as mg


def np_knn(P, Q, k, exclude_self):
    """Independent brute force: lexsort on (d2, index), FP64 from the same values."""
    out = np.zeros((len(Q), k), np.int32)
    idx = np.arange(len(P))
    for i, q in enumerate(Q):
        d = P - q
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        if exclude_self:
            d2 = d2.copy()
            d2[i] = np.inf
        o = np.lexsort((idx, d2))
        out[i] = o[:k]
    return out


# ------------------------------------------------------------------------ KNN
def test_knn_matches_independent_bruteforce(orc):
    rng = np.random.default_rng(0)
    P = orc.f32r(rng.normal(size=(700, 3)))
    for k in (1, 5, 24):
        ref = np_knn(P, P, k, True)
        assert np.array_equal(orc.knn(P, P, k, True, brute=True), ref)
        assert np.array_equal(orc.knn(P, P, k, True, brute=False), ref)
    Q = orc.f32r(rng.normal(size=(300, 3)) * 1.2)
    assert np.array_equal(orc.knn(P, Q, 5, False), np_knn(P, Q, 5, False))


def test_knn_lattice_ties_break_by_index(orc):
    g = np.arange(6, dtype=np.float64)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    perm = np.random.default_rng(1).permutation(len(P))
    P = P[perm]
    got = orc.knn(P, P, 18, True)
    assert np.array_equal(got, np_knn(P, P, 18, True))
    # every row: keys ascending, ties ascending by index, self excluded, distinct
    for i in range(len(P)):
        d2 = ((P[got[i]] - P[i]) ** 2).sum(1)
        assert np.all(np.diff(d2) >= 0)
        ties = np.diff(d2) == 0
        assert np.all(np.diff(got[i])[ties] > 0)
        assert i not in got[i] and len(set(got[i])) == 18


def test_knn_grid_equals_brute_on_surface_sites(orc):
    V, F = mg.icosphere(5)
    S = orc.f32r(mg.sample_surface(V, F, 2000, seed=0))
    assert np.array_equal(orc.knn(S, S, 20, True), orc.knn(S, S, 20, True, brute=True))
    # excluded points are never closer than the K-th
    knn = orc.knn(S, S, 20, True)
    for i in range(0, 2000, 97):
        d2 = ((S - S[i]) ** 2).sum(1)
        kth = d2[knn[i, -1]]
        mask = np.ones(len(S), bool)
        mask[knn[i]] = False
        mask[i] = False
        assert d2[mask].min() >= kth


# --------------------------------------------------------------------- frames
def test_frames_closed_forms(orc):
    U, W = orc.frames(np.array([[0.0, 0, 1], [1.0, 0, 0], [0.0, 0, 2.0]]))
    np.testing.assert_allclose(U[0], [0, -1, 0], atol=1e-15)
    np.testing.assert_allclose(W[0], [1, 0, 0], atol=1e-15)
    np.testing.assert_allclose(U[1], [0, 1, 0], atol=1e-15)
    np.testing.assert_allclose(W[1], [0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(U[2], [0, -1, 0], atol=1e-15)  # normal is normalised first


def test_frames_orthonormal_and_switch(orc):
    rng = np.random.default_rng(2)
    N = rng.normal(size=(500, 3))
    U, W = orc.frames(N)
    n = N / np.linalg.norm(N, axis=1, keepdims=True)
    np.testing.assert_allclose(np.linalg.norm(U, axis=1), 1, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(W, axis=1), 1, atol=1e-12)
    np.testing.assert_allclose((U * W).sum(1), 0, atol=1e-12)
    np.testing.assert_allclose((U * n).sum(1), 0, atol=1e-12)
    np.testing.assert_allclose((np.cross(U, W) * n).sum(1), 1, atol=1e-12)
    # exactly |n_z| = 0.9 is not "> 0.9": the helper axis stays z, so u has no z part
    nz = np.array([[np.sqrt(1 - 0.81), 0.0, 0.9]])
    nz = nz / np.linalg.norm(nz)
    if abs(nz[0, 2]) <= 0.9:
        u, _ = orc.frames(nz)
        assert abs(u[0, 2]) < 1e-15


# ---------------------------------------------------------------------- cells
def _planar(S2):
    S = np.zeros((len(S2), 3))
    S[:, :2] = S2
    return S


def _cell(orc, S, i, nb, R=10.0):
    # frame for n = +z: u = (0,-1,0), v = (1,0,0)
    u = np.array([0.0, -1.0, 0.0])
    v = np.array([1.0, 0.0, 0.0])
    return orc.cell_poly(S, i, u, v, np.asarray(nb, np.int32), R)


def test_square_lattice_cell(orc):
    a = 0.3
    g = np.arange(-3, 4) * a
    S = _planar(np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2))
    i = int(np.argmin(np.linalg.norm(S, axis=1)))
    nb = np_knn(S, S, 8, True)[i]
    poly, cent, area, mom = _cell(orc, S, i, nb)
    np.testing.assert_allclose(cent, S[i], atol=1e-14)
    assert area == pytest.approx(a * a, rel=1e-12)
    assert mom == pytest.approx(a ** 4 / 6, rel=1e-12)


def test_hex_lattice_cell(orc):
    a = 0.2
    pts = []
    for r in range(-4, 5):
        for c in range(-4, 5):
            pts.append((c * a + (r % 2) * a / 2, r * a * np.sqrt(3) / 2))
    S = _planar(np.array(pts))
    i = int(np.argmin(np.linalg.norm(S - S.mean(0), axis=1)))
    nb = np_knn(S, S, 12, True)[i]
    _, cent, area, mom = _cell(orc, S, i, nb)
    s = a / np.sqrt(3)  # hexagon side
    np.testing.assert_allclose(cent, S[i], atol=1e-13)
    assert area == pytest.approx(np.sqrt(3) / 2 * a * a, rel=1e-12)
    assert mom == pytest.approx(5 * np.sqrt(3) / 8 * s ** 4, rel=1e-12)


def test_single_neighbour_and_empty_list(orc):
    R = 1.0
    d = 0.4
    S = _planar(np.array([[0.0, 0.0], [0.0, d]]))  # neighbour along +y = +v direction
    _, cent, area, _ = _cell(orc, S, 0, [1], R)
    # half-plane v <= d/2 inside the square: rectangle [-R, R] x [-R, d/2] in (u, v)
    assert area == pytest.approx(2 * R * (R + d / 2), rel=1e-12)
    np.testing.assert_allclose(cent, [0.0, (d / 2 - R) / 2, 0.0], atol=1e-14)
    _, cent, area, mom = _cell(orc, S, 0, np.zeros(0, np.int32), R)
    assert area == pytest.approx(4 * R * R)
    np.testing.assert_allclose(cent, [0, 0, 0], atol=1e-15)
    assert mom == pytest.approx((2 * R) ** 4 / 6)


def test_duplicate_neighbour_does_not_clip(orc):
    S = _planar(np.array([[0.0, 0.0], [0.0, 0.0], [0.5, 0.0]]))
    _, c1, a1, _ = _cell(orc, S, 0, [1, 2], 1.0)
    _, c2, a2, _ = _cell(orc, S, 0, [2], 1.0)
    np.testing.assert_allclose(c1, c2)
    assert a1 == a2


def test_icosahedron_vertices_are_fixed_points(orc):
    V, F = mg.icosphere(4)
    ico, _ = mg.icosphere(0)
    mesh = orc.Mesh(V, F)
    st = orc.LloydState(ico, 5)
    out = orc.lloyd_step(mesh, st, K=5, R=np.float32(mesh.bbox_max_edge))
    np.testing.assert_allclose(out["Snew"], orc.f32r(ico), atol=1e-6)
    # each cell is a regular pentagon: all 12 areas equal
    np.testing.assert_allclose(out["area"], out["area"][0], rtol=1e-6)


# -------------------------------------------------------------- closest point
def _closest_independent(P, A, B, C):
    cands = [A, B, C]
    for X, Y in ((A, B), (B, C), (C, A)):
        t = np.clip(np.dot(P - X, Y - X) / np.dot(Y - X, Y - X), 0, 1)
        cands.append(X + t * (Y - X))
    n = np.cross(B - A, C - A)
    Q = P - np.dot(P - A, n) / np.dot(n, n) * n
    # inside test by barycentric signs
    w = [np.dot(np.cross(Y - X, Q - X), n) for X, Y in ((A, B), (B, C), (C, A))]
    if min(w) >= 0:
        cands.append(Q)
    d = [np.sum((c - P) ** 2) for c in cands]
    return cands[int(np.argmin(d))], min(d)


def test_closest_point_tri_all_regions(orc):
    rng = np.random.default_rng(3)
    A, B, C = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.2, 0.9, 0.1])
    for _ in range(2000):
        P = rng.normal(size=3) * 1.5 + np.array([0.4, 0.3, 0])
        got = orc.closest_point_tri(P, A, B, C)
        ref, d2 = _closest_independent(P, A, B, C)
        assert np.sum((got - P) ** 2) == pytest.approx(d2, rel=1e-9, abs=1e-12)
        np.testing.assert_allclose(got, ref, atol=1e-7)


def test_projection_onto_icosphere_is_radial(orc):
    V, F = mg.icosphere(5)
    mesh = orc.Mesh(V, F)
    rng = np.random.default_rng(4)
    D = rng.normal(size=(300, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    Cpts = orc.f32r(D * (1 + rng.uniform(-0.01, 0.01, size=(300, 1))))
    knnv = orc.knn(mesh.V, Cpts, 5, False)
    P = orc.project(Cpts, knnv, mesh.V, mesh.F, mesh.vf)
    # on the polyhedron: radius within the chord sag of a level-5 icosphere
    r = np.linalg.norm(P, axis=1)
    assert r.max() <= 1 + 1e-9 and r.min() > 0.998
    cosang = (P * D).sum(1) / r
    assert cosang.min() > 1 - 1e-5
    # local 5-vertex result never beats the global minimum over all faces
    gd = orc.point_mesh_d2(Cpts, mesh.V, mesh.F, use_bvh=False)
    ld = ((P - Cpts) ** 2).sum(1)
    assert np.all(ld >= gd - 1e-15)
    np.testing.assert_allclose(ld, gd, rtol=1e-9, atol=1e-15)


# -------------------------------------------------------------------- normals
def test_vertex_normals(orc):
    V, F = mg.icosphere(4)
    N = orc.vertex_normals(V, F)
    r = V / np.linalg.norm(V, axis=1, keepdims=True)
    assert np.abs(N - r).max() < 1e-2  # area weighting on irregular valence: O(edge^2)
    g = np.arange(5, dtype=np.float64)
    xs, ys = np.meshgrid(g, g, indexing="ij")
    Vp = np.stack([xs.ravel(), ys.ravel(), np.zeros(25)], 1)
    idx = np.arange(25).reshape(5, 5)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    Fp = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    Np = orc.vertex_normals(Vp, Fp)
    np.testing.assert_array_equal(Np, np.tile([0.0, 0.0, 1.0], (25, 1)))


# --------------------------------------------------------------------- freeze
def _fz(orc, same_seq, low_seq, need, thr2=1.0):
    """Drive one site through a sequence of (same, low) transitions."""
    K = 2
    st = np.zeros(1, np.uint8)
    fz = np.zeros(1, np.uint8)
    hist = []
    for same, low in zip(same_seq, low_seq):
        knn = np.array([[1, 2]], np.int32)
        prev = knn if same else np.array([[2, 1]], np.int32)
        S = np.zeros((1, 3))
        Snew = np.array([[np.sqrt(thr2) * (1.0 if low else 1.5), 0, 0]])
        st, fz, counts = orc.freeze(S, Snew, knn, prev, None, np.full(6, thr2), np.full(6, need, np.int32),
                                    1, 1, st, fz)
        hist.append((int(st[0]), int(fz[0]), tuple(counts)))
    return hist


def test_freeze_streak_scenarios(orc):
    # need = 3 -> transitions needed = 2: frozen at the second consecutive success
    h = _fz(orc, [1, 1, 1], [1, 1, 1], need=3)
    assert [x[:2] for x in h] == [(1, 0), (2, 1), (3, 1)]
    assert h[1][2][3] == 1 and h[2][2][3] == 0  # newly frozen counted once
    # a failure resets the streak; frozen is absorbing
    h = _fz(orc, [1, 0, 1, 1, 0], [1, 1, 1, 1, 1], need=3)
    assert [x[:2] for x in h] == [(1, 0), (0, 0), (1, 0), (2, 1), (0, 1)]
    # displacement exactly at the threshold counts as low (<=)
    h = _fz(orc, [1], [1], need=1)
    assert h[0][2][1] == 1
    # need <= 2 -> at least one transition
    h = _fz(orc, [1], [1], need=0)
    assert h[0][:2] == (1, 1)


def test_freeze_saturation_and_no_prev(orc):
    S = np.zeros((1, 3))
    knn = np.array([[3]], np.int32)
    st, fz, _ = orc.freeze(S, S, knn, knn, None, np.ones(6), np.full(6, 300, np.int32), 1, 1,
                           np.array([255], np.uint8), np.zeros(1, np.uint8))
    assert st[0] == 255
    st, fz, counts = orc.freeze(S, S, knn, knn, None, np.ones(6), np.full(6, 2, np.int32), 0, 1,
                                np.array([7], np.uint8), np.zeros(1, np.uint8))
    assert st[0] == 0 and fz[0] == 0 and counts.sum() == 0
    # tier ids above 5 clamp to tier 5
    st, fz, _ = orc.freeze(S, S + 0.5, knn, knn, np.array([9], np.uint8),
                           np.array([0, 0, 0, 0, 0, 1.0]), np.full(6, 9, np.int32), 1, 1,
                           np.zeros(1, np.uint8), np.zeros(1, np.uint8))
    assert st[0] == 1


# ------------------------------------------------------------------ Hausdorff
def test_point_mesh_distance(orc):
    V, F = mg.torus(1.0, 0.3, 40, 16)
    rng = np.random.default_rng(5)
    P = rng.uniform(-1.5, 1.5, size=(400, 3))
    a = orc.point_mesh_d2(P, V, F, use_bvh=False)
    b = orc.point_mesh_d2(P, V, F, use_bvh=True)
    np.testing.assert_array_equal(a, b)
    assert orc.point_mesh_d2(V, V, F).max() == 0.0
    # translated plane: every sample is exactly t away
    g = np.linspace(0, 1, 9)
    xs, ys = np.meshgrid(g, g, indexing="ij")
    Vp = np.stack([xs.ravel(), ys.ravel(), np.zeros(81)], 1)
    idx = np.arange(81).reshape(9, 9)
    a_, b_, c_, d_ = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    Fp = np.concatenate([np.stack([a_, b_, c_], 1), np.stack([a_, c_, d_], 1)])
    Q = mg.sample_surface(Vp, Fp, 200, seed=1) + np.array([0, 0, 0.125])
    np.testing.assert_allclose(orc.point_mesh_d2(Q, Vp, Fp), 0.125 ** 2, rtol=1e-12)


# -------------------------------------------------------------------- full step
def test_lloyd_energy_decreases_and_frozen_lists_pinned(orc):
    V, F = mg.icosphere(4)
    mesh = orc.Mesh(V, F)
    S = mg.sample_surface(V, F, 800, seed=3)
    st = orc.LloydState(S, 12)
    R = np.float32(mesh.bbox_max_edge)
    E = []
    for it in range(6):
        out = orc.lloyd_step(mesh, st, K=12, R=R)
        E.append(out["moment"].sum())
        st.S = orc.f32r(out["Snew"])
    assert all(E[i + 1] < E[i] for i in range(len(E) - 1))
    # frozen sites keep their lists
    st.frozen[::4] = 1
    knn0, knnv0 = st.knn.copy(), st.knnV.copy()
    orc.lloyd_step(mesh, st, K=12, R=R, freeze_mode=1, thresh2_tier=np.zeros(6), streak_tier=np.full(6, 5, np.int32))
    assert np.array_equal(st.knn[::4], knn0[::4])
    assert np.array_equal(st.knnV[::4], knnv0[::4])


def test_lloyd_exclude_self_lists(orc):
    """Sites at the mesh vertices (the command line's start), exclude_self:
    the centroid -> mesh lists are the independent brute force with vertex i
    removed for site i (main.cu:728, near_project.cuh:72-74), so no list holds
    its own site's vertex; without the switch most lists start with it."""
    V, F = mg.icosphere(3)
    mesh = orc.Mesh(V, F)
    S = mesh.V + 1e-4 * np.random.default_rng(5).standard_normal(mesh.V.shape)
    R = np.float32(mesh.bbox_max_edge)
    st = orc.LloydState(S, 12)
    out = orc.lloyd_step(mesh, st, K=12, R=R, exclude_self=True)
    assert np.array_equal(st.knnV, np_knn(mesh.V, out["cent"], 5, True))
    assert not (st.knnV == np.arange(len(S))[:, None]).any()
    st0 = orc.LloydState(S, 12)
    orc.lloyd_step(mesh, st0, K=12, R=R)
    assert np.array_equal(st0.knnV, np_knn(mesh.V, out["cent"], 5, False))
    assert (st0.knnV[:, 0] == np.arange(len(S))).mean() > 0.9


def test_lloyd_kproj10_lists_and_projection(orc):
    """kproj = 10 (scripts/testfreeze.py:379): the lists are the independent
    brute force, and on a convex mesh the projection through the 10 nearest
    vertices' faces is the global nearest point (distance from a numpy brute
    force over every face)."""
    V, F = mg.icosphere(3)
    mesh = orc.Mesh(V, F)
    S = mg.sample_surface(V, F, 300, seed=8)
    st = orc.LloydState(S, 12, kproj=10)
    out = orc.lloyd_step(mesh, st, K=12, kproj=10, R=np.float32(mesh.bbox_max_edge))
    C = out["cent"]
    assert np.array_equal(st.knnV, np_knn(mesh.V, C, 10, False))
    d_proj = np.linalg.norm(out["Snew"] - C, axis=1)
    d_all = np.sqrt(np_point_tri_min(C, mesh.V, mesh.F))
    np.testing.assert_allclose(d_proj, d_all, rtol=1e-9, atol=1e-12)


def np_point_tri_min(Q, V, F):
    """Squared distance from each query to the nearest of all triangles, by
    sampling-free brute force: the plane foot when it falls inside (barycentric
    test), else the nearest point on the three edges (independent of the
    oracle's region code)."""
    A, B, Cc = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    out = np.empty(len(Q))
    for i, q in enumerate(Q):
        n = np.cross(B - A, Cc - A)
        nn = np.einsum("ij,ij->i", n, n)
        t = np.einsum("ij,ij->i", q - A, n) / nn
        foot = q - t[:, None] * n
        def bary_in(P):
            v0, v1, v2 = B - A, Cc - A, P - A
            d00 = np.einsum("ij,ij->i", v0, v0); d01 = np.einsum("ij,ij->i", v0, v1)
            d11 = np.einsum("ij,ij->i", v1, v1); d20 = np.einsum("ij,ij->i", v2, v0)
            d21 = np.einsum("ij,ij->i", v2, v1)
            den = d00 * d11 - d01 * d01
            v = (d11 * d20 - d01 * d21) / den
            w = (d00 * d21 - d01 * d20) / den
            return (v >= 0) & (w >= 0) & (v + w <= 1)
        best = np.where(bary_in(foot), np.einsum("ij,ij->i", q - foot, q - foot), np.inf)
        for P0, P1 in ((A, B), (B, Cc), (Cc, A)):
            e = P1 - P0
            u = np.clip(np.einsum("ij,ij->i", q - P0, e) / np.einsum("ij,ij->i", e, e), 0.0, 1.0)
            X = P0 + u[:, None] * e
            best = np.minimum(best, np.einsum("ij,ij->i", q - X, q - X))
        out[i] = best.min()
    return out


# --------------------------------------------------------------------- remesh
def test_remesh_planar_points_give_delaunay(orc):
    """Dual triangles of planar Voronoi cells = the Delaunay triangulation
    (scipy/Qhull, independent) on triangles away from the boundary."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(11)
    P2 = rng.random((600, 2))
    S = _planar(P2)
    n = len(S)
    knn = np_knn(S, S, 40, True)  # enough neighbours that every true Voronoi neighbour is listed
    U = np.tile([0.0, -1.0, 0.0], (n, 1))
    W = np.tile([1.0, 0.0, 0.0], (n, 1))
    N = np.tile([0.0, 0.0, 1.0], (n, 1))
    T = orc.remesh(S, U, W, knn, 10.0, N, eps=1e-12)
    ref = np.sort(Delaunay(P2).simplices, axis=1)
    inner = lambda tri: np.all((P2[tri] > 0.15) & (P2[tri] < 0.85), axis=(1, 2))  # noqa: E731
    got = {tuple(t) for t in np.sort(T, axis=1)[inner(T)]}
    exp = {tuple(t) for t in ref[inner(ref)]}
    assert got == exp and len(exp) > 300
    # orientation: every triangle counter-clockwise seen from +z
    a, b, c = S[T[:, 0]], S[T[:, 1]], S[T[:, 2]]
    assert np.all(np.cross(b - a, c - a)[:, 2] > 0)


def test_remesh_sphere_is_closed_manifold(orc):
    V, F = mg.icosphere(4)
    mesh = orc.Mesh(V, F)
    S = orc.f32r(mg.sample_surface(V, F, 2000, seed=1))
    st = orc.LloydState(S, 12)
    for _ in range(4):
        out = orc.lloyd_step(mesh, st, K=12, R=np.float32(mesh.bbox_max_edge))
        st.S = orc.f32r(out["Snew"])
    S = st.S
    knn = orc.knn(S, S, 12, True)
    nn = orc.knn(mesh.V, S, 1, False)[:, 0]
    U, W = orc.frames(mesh.N[nn])
    T = orc.remesh(S, U, W, knn, np.float32(mesh.bbox_max_edge), mesh.N[nn])
    E = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(E, axis=0, return_counts=True)
    # tangent-plane cells need not agree across neighbours, so a few edges may
    # be shared by 1 or 3 triangles; the surface is still almost a closed sphere
    assert np.mean(cnt == 2) > 0.995
    assert abs(len(T) - (2 * len(S) - 4)) <= 0.01 * len(T)
    a, b, c = S[T[:, 0]], S[T[:, 1]], S[T[:, 2]]
    assert np.all((np.cross(b - a, c - a) * (a + b + c)).sum(1) > 0)  # outward


# ---------------------------------------------------------------------- tiers
def _tier_case(normal0, neighbour_normals):
    """Site 0 with the given normal and neighbours 1..K with theirs (sites are
    their own vertices); returns the oracle's (tier, nv, L) of site 0 for
    schemes 1 and 2."""
    K = len(neighbour_normals)
    N = np.vstack([normal0, neighbour_normals]).astype(np.float64)
    n = K + 1
    knn = np.zeros((n, K), np.int32)
    knn[0] = np.arange(1, K + 1)
    for i in range(1, n):
        knn[i] = [j for j in range(n) if j != i][:K]
    nn = np.arange(n, dtype=np.int32)
    return [orc_tiers_site0(knn, nn, N, s) for s in (1, 2)]


def orc_tiers_site0(knn, nn, N, scheme):
    from oracle import oracle as o
    t, nv, L = o.tiers(knn, nn, N, scheme)
    return int(t[0]), float(nv[0]), float(L[0])


def test_tiers_plane_is_tier0(orc):
    K = 16
    (t1, nv1, L1), (t2, nv2, L2) = _tier_case([0, 0, 1], np.tile([0.0, 0.0, 1.0], (K, 1)))
    assert (t1, t2) == (0, 0) and nv1 == 0.0 and L2 == 0.0


@pytest.mark.parametrize("k2, theta_deg", [(4, 30.0), (8, 60.0), (8, 90.0), (12, 100.0), (8, 150.0), (16, 120.0)])
def test_tiers_crease_closed_form(orc, k2, theta_deg):
    """k1 = K - k2 neighbours share the site's normal a, k2 have normal b at
    angle theta: nv = (k2/K)(1 - cos theta); the normals form two clusters, so
    their covariance has rank one and L = 1 (up to rounding)."""
    K = 16
    th = np.radians(theta_deg)
    a = np.array([0.0, 0.0, 1.0])
    b = np.array([np.sin(th), 0.0, np.cos(th)])
    nb = np.vstack([np.tile(a, (K - k2, 1)), np.tile(b, (k2, 1))])
    (t1, nv1, _), (t2, nv2, L2) = _tier_case(a, nb)
    nv_exp = (k2 / K) * (1.0 - np.cos(th))
    assert abs(nv1 - nv_exp) < 1e-12 and abs(nv2 - nv_exp) < 1e-12
    tier_exp = int(np.sum(nv_exp >= np.array([0.15, 0.35, 0.55, 0.80])))
    assert t1 == tier_exp
    if k2 < K:
        assert abs(L2 - 1.0) < 1e-9
        assert t2 == (5 if tier_exp == 4 else tier_exp)
    else:  # one cluster only: zero covariance, L = 0
        assert L2 == 0.0 and t2 == tier_exp


def test_tiers_cone_has_no_split(orc):
    """Neighbour normals evenly spaced on a cone of half-angle 85 deg about the
    site's normal: nv = 1 - cos 85 deg (tier 4) and the covariance has two equal
    largest eigenvalues, so L ~ 0 and scheme 2 keeps tier 4 (a crease gives 5)."""
    K = 16
    al = np.radians(85.0)
    phi = 2 * np.pi * np.arange(K) / K
    nb = np.stack([np.sin(al) * np.cos(phi), np.sin(al) * np.sin(phi), np.full(K, np.cos(al))], axis=1)
    (t1, nv1, _), (t2, nv2, L2) = _tier_case([0.0, 0.0, 1.0], nb)
    assert abs(nv1 - (1 - np.cos(al))) < 1e-12
    assert t1 == 4 and t2 == 4 and L2 < 1e-6


# -------------------------------------------------------------------- quality
def _iso(apex_deg):
    h = np.radians(apex_deg) / 2
    return np.array([[0.0, 0.0, 0.0], [np.sin(h), -np.cos(h), 0.0], [-np.sin(h), -np.cos(h), 0.0]])


def test_quality_closed_forms(orc):
    eq = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0]])
    q = orc.quality(eq, [[0, 1, 2]])
    assert abs(q["Qmin"] - 1) < 1e-6 and abs(q["theta_min"] - 60) < 1e-4
    assert q["theta_lt_30_pct"] == 0 and q["theta_gt_90_pct"] == 0
    ri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    q = orc.quality(ri, [[0, 1, 2]])
    assert abs(q["Qmin"] - np.sqrt(3) * (np.sqrt(2) - 1)) < 1e-6 and abs(q["theta_min"] - 45) < 1e-4
    # apex 130 degrees: base angles 25, legs 1, base 2 sin 65
    q = orc.quality(_iso(130.0), [[0, 1, 2]])
    s65 = np.sin(np.radians(65.0))
    q_cf = (6 / np.sqrt(3)) * (0.5 * np.sin(np.radians(130.0))) / ((1 + s65) * 2 * s65)
    assert abs(q["Qmin"] - q_cf) < 1e-6 and abs(q["theta_min"] - 25) < 1e-4
    assert abs(q["theta_lt_30_pct"] - 200 / 3) < 1e-4 and abs(q["theta_gt_90_pct"] - 100 / 3) < 1e-4
    # collinear: no area, angles 0, 0, 180
    q = orc.quality(np.array([[0.0, 0, 0], [1.0, 0, 0], [3.0, 0, 0]]), [[0, 1, 2]])
    assert q["Qmin"] == 0 and q["theta_min"] == 0
    assert abs(q["theta_lt_30_pct"] - 200 / 3) < 1e-4 and abs(q["theta_gt_90_pct"] - 100 / 3) < 1e-4


def test_quality_means_and_invariance(orc):
    P = np.concatenate([np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0]]),
                        np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]) + 5.0])
    q = orc.quality(P, [[0, 1, 2], [3, 4, 5]])
    assert abs(q["Qavg"] - (1 + np.sqrt(3) * (np.sqrt(2) - 1)) / 2) < 1e-6
    assert abs(q["theta_min_avg"] - (60 + 45) / 2) < 1e-4
    assert abs(q["Qmin"] - np.sqrt(3) * (np.sqrt(2) - 1)) < 1e-6
    V, F = mg.displaced_sphere(10, seed=2, crease_seed=3)
    rng = np.random.default_rng(4)
    Rm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    q0, q1 = orc.quality(V, F), orc.quality(3.7 * V @ Rm.T + 1.0, F)
    for k in ("Qmin", "Qavg", "theta_min", "theta_min_avg"):
        assert abs(q0[k] - q1[k]) <= 2e-5 * max(1.0, abs(q0[k])), k
