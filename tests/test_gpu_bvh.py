"""GPU LBVH point-to-mesh distances against the oracle (brute force / BVH)."""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", ["torus", "ico", "dsphere", "single"])
def test_point_mesh_d2_matches_oracle(orc, mesh):
    if mesh == "torus":
        V, F = mg.torus(1.0, 0.3, 80, 30, jitter=1e-3, seed=1)
    elif mesh == "ico":
        V, F = mg.icosphere(4)
    elif mesh == "dsphere":
        V, F = mg.displaced_sphere(24, seed=4, crease_seed=5)
    else:
        V = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]])
        F = np.array([[0, 1, 2]], np.int32)
    print(f"[bvh] mesh={mesh} nF={len(F)}", flush=True)
    rng = np.random.default_rng(7)
    P = np.concatenate([rng.uniform(-1.6, 1.6, size=(3000, 3)), mg.sample_surface(V, F, 1000, seed=3)])
    Vr = orc.f32r(V)
    P32 = orc.f32r(P)
    ref = orc.point_mesh_d2(P32, Vr, F, use_bvh=True)
    b = cvt.BVH(V, F)
    got = b.point_d2(P)
    b.close()
    # FP32 traversal: squared distances agree to FP32 rounding of the closest-point formula
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-10)


def test_job_point_mesh_and_hausdorff(orc):
    V, F = mg.icosphere(4)
    S = mg.sample_surface(V, F, 500, seed=2)
    with cvt.Job(V, F, S, K=8) as job:
        d2 = job.point_mesh_d2(S)
    assert d2.max() < 1e-12
    V2, F2 = mg.icosphere(3)
    h = cvt.hausdorff(V2, F2, V, F, 20000, seed=5)
    # coarse icosphere lies inside the fine one: the deviation is bounded by the
    # level-3 chord sag and is far above zero
    Vr, Vc = orc.f32r(V), orc.f32r(V2)
    Pa = mg.sample_surface(V2, F2, 20000, 5)
    Pb = mg.sample_surface(V, F, 20000, 6)
    ref = max(np.sqrt(orc.point_mesh_d2(orc.f32r(Pa), Vr, F).max()),
              np.sqrt(orc.point_mesh_d2(orc.f32r(Pb), Vc, F2).max()))
    assert h == pytest.approx(ref, rel=1e-4)
    assert 1e-3 < h < 0.02
