"""The allocation-free LBVH query (cvt_bvh_point_d2_noalloc) through the C ABI
against the allocating one (cvt_bvh_point_d2): bit-identical distances on two
calls in a row, the second with fewer points (the shared device staging area
keeps the first call's contents), plus the size limit and the empty call."""
import ctypes as C

import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu

MAX_POINTS = 1048576  # CVT_NOALLOC_MAX_POINTS in include/cvt.h


def _query(fn, h, P):
    out = np.full(len(P), np.nan)
    rc = fn(h, np.ascontiguousarray(P, np.float64).reshape(-1), len(P), out)
    return rc, out


def test_noalloc_matches_allocating_query():
    L = cvt.lib()
    V, F = mg.torus(1.0, 0.3, 120, 48, jitter=1e-3, seed=3)
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    h = C.c_void_p()
    assert L.cvt_bvh_create(V.reshape(-1), len(V), F.reshape(-1), len(F), C.byref(h)) == 0
    try:
        rng = np.random.default_rng(5)
        for n in (200_000, 3_001):  # the second call is smaller
            P = rng.uniform(-1.6, 1.6, size=(n, 3))
            print(f"[noalloc] n={n}", flush=True)
            rc_a, ref = _query(L.cvt_bvh_point_d2, h, P)
            rc_b, got = _query(L.cvt_bvh_point_d2_noalloc, h, P)
            assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, L.cvt_last_error())
            assert np.array_equal(got, ref), f"n={n}: {np.sum(got != ref)} distances differ"
            assert np.all(np.isfinite(got)) and np.all(got >= 0)
        # empty call: nothing staged, nothing written
        rc, _ = _query(L.cvt_bvh_point_d2_noalloc, h, np.zeros((0, 3)))
        assert rc == 0
        # beyond the staging area: a clear error, no device work
        big = np.zeros(3 * (MAX_POINTS + 1))
        out = np.zeros(MAX_POINTS + 1)
        rc = L.cvt_bvh_point_d2_noalloc(h, big, MAX_POINTS + 1, out)
        assert rc == -1  # CVT_EINVAL
        assert b"exceed" in L.cvt_last_error()
    finally:
        L.cvt_bvh_destroy(h)


def test_noalloc_matches_oracle_distances(orc):
    """Small mesh: the staged query agrees with the oracle's brute-force
    point-mesh distance (FP32 geometry, so within FP32 rounding)."""
    V, F = mg.icosphere(2)
    P = np.random.default_rng(9).normal(size=(500, 3))
    b = cvt.BVH(V, F)
    try:
        got = b.point_d2_noalloc(P)
    finally:
        b.close()
    ref = orc.point_mesh_d2(orc.f32r(P), orc.f32r(V), F, use_bvh=False)
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-9)

