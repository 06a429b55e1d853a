"""Triangle quality on the device (eval_quality_angles_cpu,
evaluate.cuh:388-462) against the oracle's FP32 restatement: the shape
quality uses only correctly rounded operations, so its minimum is
bit-identical; angles go through acosf, whose device and host versions may
differ in the last place, so angle statistics agree to a few ulps and the
30/90-degree shares to a borderline angle or two."""
import numpy as np
import pytest

from gpu_cvt_amd import cvt, meshgen as mg

pytestmark = pytest.mark.gpu


def _compare(got, ref, nF):
    print(f"  device {got}\n  oracle {ref}", flush=True)
    assert got["nF"] == nF
    assert got["Qmin"] == ref["Qmin"]
    assert abs(got["Qavg"] - ref["Qavg"]) <= 1e-6 * ref["Qavg"]
    for k in ("theta_min", "theta_min_avg"):
        assert abs(got[k] - ref[k]) <= 2e-6 * max(ref[k], 1.0), k
    for k in ("theta_lt_30_pct", "theta_gt_90_pct"):
        assert abs(got[k] - ref[k]) <= 2 * 100.0 / (3 * nF) + 1e-5, k


@pytest.mark.parametrize("shape", ["dsphere", "torus"])
def test_mesh_quality_matches_oracle(orc, shape):
    V, F = mg.displaced_sphere(60, seed=61, crease_seed=62) if shape == "dsphere" else \
        mg.torus(1.0, 0.3, 120, 48, jitter=2e-3, seed=63)
    print(f"[quality] {shape} nV={len(V)} nF={len(F)}", flush=True)
    _compare(cvt.mesh_quality(V, F), orc.quality(V, F), len(F))


def test_job_quality_of_remesh_matches_oracle(orc):
    V, F = mg.displaced_sphere(40, seed=64, crease_seed=65)
    S = mg.sample_surface(V, F, 5000, seed=66)
    with cvt.Job(V, F, S, K=24) as job:
        job.step(6)
        T = job.remesh()
        got = job.quality(T)
        Sf = job.sites()
    print(f"[quality] remesh of {len(S)} sites: {len(T)} triangles", flush=True)
    _compare(got, orc.quality(Sf, T), len(T))


def test_quality_rejects_bad_ids():
    V, F = mg.icosphere(2)
    bad = F.copy()
    bad[3, 1] = len(V)
    with pytest.raises(cvt.CVTError):
        cvt.mesh_quality(V, bad)
