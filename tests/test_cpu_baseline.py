"""The FP32 CPU Lloyd step timed as the CPU baseline computes the same
iteration as the FP64 oracle: one step from the same FP32 sites agrees within
FP32 rounding (1e-5 of the bbox diagonal), and several steps keep agreeing."""
import numpy as np

from gpu_cvt_amd import meshgen as mg


def test_fp32_baseline_step_matches_oracle(orc):
    V, F = mg.icosphere(4)
    S = mg.sample_surface(V, F, 2000, seed=3)
    mesh = orc.Mesh(V, F)
    R = float(np.float32(mesh.bbox_max_edge))
    for K in (8, 20):
        st = orc.LloydState(S, K)
        ref = orc.lloyd_step(mesh, st, K=K, R=R)["Snew"]
        got = orc.cpu32_lloyd(mesh, orc.f32r(S), K, R, 1, 2)
        err = np.linalg.norm(got - ref, axis=1).max()
        assert err <= 1e-5 * mesh.bbox_diag, f"K={K}: {err / mesh.bbox_diag:.2e} x diag"


def test_fp32_baseline_threads_agree(orc):
    V, F = mg.torus(1.0, 0.3, 60, 24, jitter=1e-3, seed=3)
    S = mg.sample_surface(V, F, 3000, seed=4)
    mesh = orc.Mesh(V, F)
    R = float(np.float32(mesh.bbox_max_edge))
    a = orc.cpu32_lloyd(mesh, S, 12, R, 3, 1)
    b = orc.cpu32_lloyd(mesh, S, 12, R, 3, 4)
    assert np.array_equal(a, b)
